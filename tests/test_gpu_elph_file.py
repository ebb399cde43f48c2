"""GPU: the coupling read from Phoebe's el-ph file (streamed from the file into
HBM) drives cacheElPh / calcCouplingSquared and the linewidth builder exactly
as the same tensor handed over from host memory, and the real Si coupling of
the reference's fixture matches the FP64 oracle."""
import os
import sys

import numpy as np
import pytest

import phoebe_amd as P
from oracle import h5_oracle as H
from oracle import oracle as O
from phoebe_amd import synth
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "silicon.phoebe.elph.hdf5")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _unitary(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_si_fixture_coupling_vs_oracle(precision, tol):
    """Real Si gW (8 Wannier functions, 6 modes, 19 + 19 lattice vectors):
    the interaction parsed from the file against the oracle on the same
    tensor, with random unitary rotations and Cartesian wavevectors."""
    say("Si fixture coupling", precision)
    o = H.read_elph(FIXTURE)
    nw, nm = o["nw"], o["nm"]
    rng = np.random.default_rng(17)
    a = 10.2
    recip = 2 * np.pi / a * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], dtype=float).T
    k1 = recip @ rng.uniform(0, 1, 3)
    U1 = _unitary(rng, nw)[:, 2:6]
    inter = P.InteractionElPhWan.parse(FIXTURE, precision=precision)
    assert list(inter.getCouplingDimensions()) == [o["nRe"], o["nRp"], nm, nw, nw]
    host = P.InteractionElPhWan(o["gW"], o["elR"], o["elDeg"], o["phR"], o["phDeg"], precision)
    nq = 23
    q3s = [recip @ rng.uniform(-1, 1, 3) for _ in range(nq)]
    U2s = [_unitary(rng, nw)[:, : 1 + j % nw] for j in range(nq)]
    e3s = [_unitary(rng, nm) for _ in range(nq)]
    for it in (inter, host):
        it.cacheElPh(U1, k1)
        it.calcCouplingSquared(U1, U2s, e3s, q3s)
    cached = O.cache_elph(nw, nm, o["phR"], o["elR"], o["elDeg"], o["gW"], U1, k1)
    ref = O.coupling_sq(nw, nm, o["phR"], o["phDeg"], cached, U2s, e3s, q3s)
    worst = 0.0
    for j in range(nq):
        got = inter.getCouplingSquared(j)
        # the file path puts the same tensor on the device as the host path
        assert np.array_equal(got, host.getCouplingSquared(j)), j
        worst = max(worst, np.abs(got - ref[j]).max() / np.abs(ref[j]).max())
    say("  max |g|^2 error / max |g|^2 =", worst)
    assert worst < tol


@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_linewidths_from_file_match_host_gw(tmp_path, fmt, precision):
    say("linewidths from elph file", fmt, precision)
    p = synth.make_problem(2, 5, 3, 3, 4, seed=61, name="file")
    gW = p.gw_host()
    path = tmp_path / "synth.phoebe.elph.hdf5"
    P.write_elph_file(path, gW, p.Re, p.degE, p.Rp, p.degP, numElectrons=4, kMesh=p.mesh,
                      qMesh=p.mesh, fileFormat=fmt)
    a = P.ElectronLinewidths(p, gW=gW, precision=precision, transform="separable")
    b = P.ElectronLinewidths(p, precision=precision, transform="separable", elph_file=path)
    if precision == "fp64":
        assert np.array_equal(a.copy_gw(), b.copy_gw())
    a.run()
    b.run()
    la, lb = a.linewidths(), b.linewidths()
    assert np.array_equal(la, lb)
    if precision == "fp64":
        ref = O.Oracle(p, gW.reshape(-1)).linewidths()
        err = rel_err(lb, ref)
        say("  max rel err vs oracle", err.max())
        assert err.max() < 1e-10


def test_streaming_upload_over_many_chunks(tmp_path):
    """An 11.8 M-element tensor (189 MB: three 64 MiB staging chunks, the last
    one partial, both staging buffers reused) arrives on the device bit-exact."""
    say("streamed upload, 189 MB")
    p = synth.make_problem(4, 8, 5, 4, 3, seed=5, name="big")
    gW = p.gw_host()
    assert 2 * (4 << 20) < gW.size < 3 * (4 << 20)
    path = tmp_path / "big.phoebe.elph.hdf5"
    P.write_elph_file(path, gW, p.Re, p.degE, p.Rp, p.degP, numElectrons=8, fileFormat=1)
    lw = P.ElectronLinewidths(p, precision="fp64", elph_file=path)
    assert np.array_equal(lw.copy_gw(), gW.reshape(-1).view(np.float64).reshape(lw.copy_gw().shape))


def test_setup_file_rejects_mismatched_bands(tmp_path):
    p = synth.make_problem(1, 4, 2, 2, 3, seed=2, name="mismatch")
    with pytest.raises(P.PhxError, match="numElBands"):
        P.ElectronLinewidths(p, elph_file=FIXTURE)
