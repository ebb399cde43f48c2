"""The oracle's phonon linewidths from el-ph scattering (PhElScatteringMatrix::
builder, phel_scattering.cpp:43-363) against a closed form: one cosine band,
three flat modes and on-site coupling make |g|^2 = |gW|^2 / M, so the
linewidth is a sum numpy evaluates from the band formula alone (k mesh 4^3,
q mesh 3^3, so k + q leaves the k mesh)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.phel_cases import toy


@pytest.mark.parametrize("springs", [(8.0, 20.0, 45.0), (20.0, 20.0, 45.0)])
def test_phel_closed_form(springs):
    p, qm = toy(springs)
    ref = p.expected_phel(qm)
    orc = O.Oracle(p, p.gw_host().reshape(-1))
    got = orc.phel_linewidths(qm)
    assert orc.phel_pairs == 64 * 27
    assert ref.max() > 0 and np.count_nonzero(ref) > 0.5 * ref.size
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_phel_closed_form_catches_planted_mistakes():
    """The closed form separates the oracle from plausible wrong formulas:
    k2 folded onto the k mesh, the emission delta instead of absorption, a
    missing 1/T; each differs from the oracle's result."""
    p, qm = toy()
    got = O.Oracle(p, p.gw_host().reshape(-1)).phel_linewidths(qm)
    scale = np.abs(got).max()
    assert np.abs(got - p.expected_phel(qm)).max() <= 1e-12 * scale
    for wrong in (dict(fold=True), dict(sign=-1.0), dict(with_T=False)):
        assert np.abs(got - p.expected_phel(qm, **wrong)).max() > 1e-3 * scale, wrong
