"""The drop-in boundary from C++: tests/cpp/adaptor_run includes the adaptor
(phoebe_amd/csrc/phx_adaptor.hpp), links libphx_elph.so and calls it with the
reference's argument types the way ElScatteringMatrix::builder calls
InteractionElPhWan; its outputs are checked against the oracle and LAPACK."""
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import h5_oracle as H
from oracle import oracle as O
from tests.helpers import projector_diff
from tests.test_gpu_polar import polar_problem

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "adaptor_run")
FIXTURE = os.path.join(ROOT, "tests", "golden", "silicon.phoebe.elph.hdf5")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cf(a):
    """complex column-major flat"""
    return np.asfortranarray(a, dtype=np.complex128).reshape(-1, order="F")


def write_input(path, gW, elR, elDeg, phR, phDeg, U1, k1, U2s, e3s, q3s, precision, pol, nspecies, mats):
    nRe, nRp, nm, nw, _ = gW.shape
    hdr = [nw, nm, nRp, nRe, U1.shape[1], len(q3s), precision, pol is not None,
           0 if pol is None else len(pol["positions"]), nspecies]
    with open(path, "wb") as f:
        np.asarray(hdr, dtype=np.int32).tofile(f)
        np.ascontiguousarray(gW, dtype=np.complex128).tofile(f)
        for a in (elR, elDeg, phR, phDeg):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        _cf(U1).tofile(f)
        np.asarray(k1, dtype=np.float64).tofile(f)
        np.asarray([u.shape[1] for u in U2s], dtype=np.int32).tofile(f)
        for u in U2s:
            _cf(u).tofile(f)
        for e in e3s:
            _cf(e).tofile(f)
        np.asarray(q3s, dtype=np.float64).tofile(f)
        if pol is not None:
            for k in ("positions", "born", "dielectric"):
                np.ascontiguousarray(pol[k], dtype=np.float64).tofile(f)
            np.asarray(pol["recip"], dtype=np.float64).reshape(-1, order="F").tofile(f)
            np.asarray([pol["volume"]], dtype=np.float64).tofile(f)
            np.asarray(pol["coarse_mesh"], dtype=np.int32).tofile(f)
        np.asarray([mats.shape[0], mats.shape[1]], dtype=np.int32).tofile(f)
        for m in mats:
            _cf(m).tofile(f)


class Out:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, dtype, n):
        a = np.frombuffer(self.b, dtype=dtype, count=n, offset=self.o)
        self.o += a.nbytes
        return a


def run(tmp_path, args):
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    say("  ", r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0, r.stderr


def _rand_herm_batch(rng, M, N):
    X = rng.standard_normal((M, N, N)) + 1j * rng.standard_normal((M, N, N))
    return X + np.conj(np.transpose(X, (0, 2, 1)))


def test_cpp_adaptor_polar_problem_vs_oracle(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/bin/adaptor_run is not built (python -c 'import __graft_entry__ as g; g.build()')"
    p = polar_problem()
    pol = dict(p.polar, recip=p.recip)
    gW = p.gw_host()
    rng = np.random.default_rng(5)
    k1 = p.recip @ rng.uniform(0, 1, 3)
    _, U1 = O.el_diag(p, k1)
    U1 = U1[:, 1:3]
    q3s = [p.recip @ rng.uniform(-1, 1, 3) for _ in range(11)] + [np.zeros(3)]
    U2s = [O.el_diag(p, k1 + q)[1][:, : 1 + j % p.nw] for j, q in enumerate(q3s)]
    e3s = [O.ph_diag(p, q)[1] for q in q3s]
    mats = _rand_herm_batch(rng, 6, 7)
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, gW, p.Re, p.degE, p.Rp, p.degP, U1, k1, U2s, e3s, q3s, 0, pol, 2, mats)
    run(tmp_path, [inp, outp])
    o = Out(outp)
    nb1, nm = U1.shape[1], p.nm
    assert list(o.take(np.int32, 5)) == [p.nRe, p.nRp, p.nm, p.nw, p.nw]
    nb = o.take(np.int32, 3)
    maxmem = float(os.environ.get("MAXMEM", 16)) * 1e9
    used = O.accounted_bytes(p.nw, p.nm, p.nRp, p.nRe, nb1_cached=nb1)
    assert nb[0] == O.estimate_num_batches(p.nw, p.nm, p.nRp, 1000, nb1, maxmem - used)
    assert nb[1] == O.estimate_num_batches(p.nw, p.nm, p.nRp, 1000, nb1, 1.0e6) > 1
    assert nb[2] == 0
    assert o.take(np.int32, 1)[0] == 1  # two species and a dielectric tensor: polar term on
    assert o.take(np.float64, 1)[0] == used
    x = o.take(np.complex128, len(q3s) * nm).reshape(len(q3s), nm)
    xo = np.stack([O.polar_part1(pol, p.recip, q, e) for q, e in zip(q3s, e3s)])
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    cached = O.cache_elph(p.nw, p.nm, p.Rp, p.Re, p.degE, gW, U1, k1)
    ref = O.coupling_sq_polar(p.nw, p.nm, p.Rp, p.degP, cached, U2s, e3s, q3s, U1, xo)
    worst = 0.0
    for j, u in enumerate(U2s):
        got = o.take(np.float64, nb1 * u.shape[1] * nm).reshape((nb1, u.shape[1], nm), order="F")
        worst = max(worst, np.abs(got - ref[j]).max() / np.abs(ref[j]).max())
    say("  |g|^2 through the C++ adaptor: max err / max", worst)
    assert worst < 1e-10
    W = o.take(np.float64, 6 * 7).reshape(6, 7)
    V = o.take(np.complex128, 6 * 49).reshape(6, 7, 7)
    for i in range(6):
        wl, vl, _ = lapack.zheev(mats[i], compute_v=1, lower=1)
        assert np.abs(W[i] - wl).max() <= 1e-12 * max(1.0, np.abs(wl).max())
        assert projector_diff(V[i].T, vl, wl) < 1e-9
    assert o.o == len(o.b)


def test_cpp_adaptor_parse_matches_constructor_on_si(tmp_path):
    """Si fixture: the constructor with the fixture's tensor and parse(context)
    on the file give the same |g|^2, equal to the oracle's."""
    assert os.path.exists(BIN), "tests/cpp/bin/adaptor_run is not built"
    f = H.read_elph(FIXTURE)
    nw, nm = f["nw"], f["nm"]
    rng = np.random.default_rng(3)
    recip = 2 * np.pi / 10.2 * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], dtype=float).T

    def unitary(n):
        q, r = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        return q * (np.diag(r) / np.abs(np.diag(r)))
    k1 = recip @ rng.uniform(0, 1, 3)
    U1 = unitary(nw)[:, 2:6]
    q3s = [recip @ rng.uniform(-1, 1, 3) for _ in range(9)]
    U2s = [unitary(nw)[:, : 1 + j % nw] for j in range(9)]
    e3s = [unitary(nm) for _ in range(9)]
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, f["gW"], f["elR"], f["elDeg"], f["phR"], f["phDeg"], U1, k1, U2s, e3s, q3s, 0, None, 1,
                np.zeros((0, 1, 1), dtype=np.complex128))
    run(tmp_path, [inp, outp, FIXTURE])
    o = Out(outp)
    dims = list(o.take(np.int32, 5))
    o.take(np.int32, 3)
    assert o.take(np.int32, 1)[0] == 0  # one species: no polar term
    o.take(np.float64, 1)
    cached = O.cache_elph(nw, nm, f["phR"], f["elR"], f["elDeg"], f["gW"], U1, k1)
    ref = O.coupling_sq(nw, nm, f["phR"], f["phDeg"], cached, U2s, e3s, q3s)
    ctor = [o.take(np.float64, U1.shape[1] * u.shape[1] * nm) for u in U2s]
    assert list(o.take(np.int32, 5)) == dims == [f["nRe"], f["nRp"], nm, nw, nw]
    parsed = [o.take(np.float64, U1.shape[1] * u.shape[1] * nm) for u in U2s]
    assert o.o == len(o.b)
    for j, u in enumerate(U2s):
        assert np.array_equal(ctor[j], parsed[j])
        got = ctor[j].reshape((U1.shape[1], u.shape[1], nm), order="F")
        assert np.abs(got - ref[j]).max() <= 1e-10 * np.abs(ref[j]).max()
