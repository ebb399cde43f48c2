"""Precision model of the R_ph transform on real coupling data (CPU).

The real-expanded GEMM g3 = A B' of the GPU path is evaluated in FP64, in
FP32, and with split-BF16 operands (FP32 accumulation), emulated with exact
round-to-nearest-even bf16 rounding.  The measured |g|^2 errors decide which
reduced-precision forms can meet the 1e-6 linewidth gate:
  * FP32, and 6-pass split-BF16 (3-way split): weighted |g|^2 error < 1e-6;
  * 3-pass split-BF16 (hi/lo, dropping lo*lo): > 1e-6, so it is not used as
    the chosen form.
"""
import numpy as np

from oracle import oracle as O
from phoebe_amd import synth


def bf16(x):
    """Round FP32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def split(x, n):
    r = x.astype(np.float32).astype(np.float64)
    parts = []
    for _ in range(n):
        h = bf16(r)
        parts.append(h)
        r = (r - h).astype(np.float32).astype(np.float64)
    return parts


def test_bf16_rounding_is_exact_rne():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415926], dtype=np.float32)
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even
    assert r[2] == 1.0078125                     # above the tie -> up
    assert abs(r[3] + 3.140625) < 1e-12


def test_split_bf16_error_model():
    p = synth.make_problem(4, 16, 4, 4, 6, seed=3)
    gW = p.gw_host()
    rng = np.random.default_rng(1)
    k1 = p.recip @ rng.uniform(0, 1, 3)
    _, U1 = O.el_diag(p, k1)
    cached = O.cache_elph(p.nw, p.nm, p.Rp, p.Re, p.degE, gW, U1[:, 7:9], k1)
    nq = 200
    qs = (p.recip @ rng.uniform(0, 1, (3, nq))).T
    ph = np.exp(1j * qs @ p.Rp.T)
    C = cached.reshape(len(p.Rp), -1) / p.degP[:, None]
    A = np.stack([ph.real, ph.imag], -1).reshape(nq, -1)
    B = np.zeros((2 * len(p.Rp), 2 * C.shape[1]))
    B[0::2, 0::2], B[0::2, 1::2] = C.real, C.imag
    B[1::2, 0::2], B[1::2, 1::2] = -C.imag, C.real
    exact = A @ B

    def gsq(g):
        return np.abs(g[:, 0::2] + 1j * g[:, 1::2]) ** 2

    def run(n, terms):
        As, Bs = split(A, n), split(B, n)
        return sum((As[i].astype(np.float32) @ Bs[j].astype(np.float32)).astype(np.float64)
                   for i, j in terms)

    ref = gsq(exact)

    def werr(g):
        return np.abs(gsq(g) - ref).sum() / ref.sum()

    e32 = werr((A.astype(np.float32) @ B.astype(np.float32)).astype(np.float64))
    e3 = werr(run(2, [(0, 0), (0, 1), (1, 0)]))
    e6 = werr(run(3, [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]))
    print(f"weighted |g|^2 error: fp32 {e32:.2e}  bf16x3 {e3:.2e}  bf16x6 {e6:.2e}")
    assert e32 < 1e-6 and e6 < 1e-6
    assert e3 > 1e-6
    assert e3 > 5 * e6
