"""ctypes front-end of the CPU FP64 oracle (oracle/elph_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcProblem(C.Structure):
    _fields_ = [
        ("recip", C.c_double * 9), ("mesh", C.c_int * 3),
        ("nw", C.c_int), ("nRh", C.c_int), ("Rh", _dp), ("degh", _dp), ("h0R", _dp),
        ("nm", C.c_int), ("nRf", C.c_int), ("Rf", _dp), ("wf", _dp), ("mat2R", _dp),
        ("masses", _dp),
        ("nRe", C.c_int), ("nRp", C.c_int), ("Re", _dp), ("degE", _dp), ("Rp", _dp),
        ("degP", _dp), ("gW", _dp),
        ("nCalc", C.c_int), ("temperatures", _dp), ("chemPots", _dp),
        ("sigma", C.c_double), ("winMin", C.c_double), ("winMax", C.c_double),
        ("natoms", C.c_int), ("nG", C.c_int), ("atomPos", _dp), ("bornCharges", _dp),
        ("dielectric", _dp), ("gVectors", _dp), ("lrConstant", _dp),
        ("lrNorm", C.c_double), ("gMax", C.c_double),
        ("smearing", C.c_int),
        ("polarOn", C.c_int), ("polarPos", _dp), ("polarBorn", _dp), ("polarEps", _dp),
        ("polarVolume", C.c_double), ("polarCoarse", C.c_int * 3),
    ]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_ws_vectors.restype = C.c_int
        L.orc_mesh_index.restype = C.c_int
        L.orc_zheev.restype = C.c_int
        L.orc_el_diag.restype = C.c_int
        L.orc_ph_diag.restype = C.c_int
        L.orc_gaussian_delta.restype = C.c_double
        L.orc_gaussian_delta.argtypes = [C.c_double, C.c_double]
        L.orc_population.restype = C.c_double
        L.orc_population.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_phonon_cutoff.restype = C.c_double
        L.orc_setup.restype = C.c_void_p
        L.orc_setup.argtypes = [C.POINTER(OrcProblem)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_num_active.argtypes = [C.c_void_p]
        L.orc_num_states.argtypes = [C.c_void_p]
        L.orc_get_window.argtypes = [C.c_void_p, _ip, _ip, _ip, _dp]
        L.orc_get_eigvecs.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_get_phonons.argtypes = [C.c_void_p, _dp, _dp, C.c_char_p]
        L.orc_linewidth_pairs.restype = C.c_longlong
        L.orc_linewidth_pairs.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, C.c_int, _dp]
        L.orc_degeneracy_average.argtypes = [C.c_void_p, _dp]
        L.orc_average_point.argtypes = [C.c_int, _dp, _dp]
        L.orc_window.restype = C.c_int
        L.orc_window.argtypes = [C.c_int, _dp, C.c_double, C.c_double, _ip, _ip]
        L.orc_num_threads.restype = C.c_int
        L.orc_adaptive_delta.restype = C.c_double
        L.orc_adaptive_delta.argtypes = [C.c_double, _dp, _dp, _ip]
        L.orc_el_group_velocities.restype = C.c_int
        L.orc_polar_part1.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, _ip, C.c_int, _dp, _dp, _dp]
        L.orc_phel.restype = C.c_longlong
        L.orc_phel.argtypes = [C.POINTER(OrcProblem), _ip, C.c_int, _ip, C.c_int, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def cplx_view(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


# ------------------------------------------------------------------ kernels
def ws_vectors(grid, A, super_cell_factor=2):
    L = lib()
    g = (C.c_int * 3)(*grid)
    Acm = _c(np.asarray(A).T.reshape(-1))  # column-major
    n = L.orc_ws_vectors(g, super_cell_factor, _p(Acm), 0, None, None)
    if n < 0:
        raise RuntimeError("WS completeness failed")
    R = np.zeros((n, 3))
    d = np.zeros(n)
    m = L.orc_ws_vectors(g, super_cell_factor, _p(Acm), n, _p(R), _p(d))
    assert m == n
    return R, d


def mesh_index(xcrys, mesh):
    x = _c(xcrys)
    m = (C.c_int * 3)(*mesh)
    return lib().orc_mesh_index(_p(x), m)


def zheev(H):
    """Eigenvalues (ascending) and eigenvectors (columns) of Hermitian H."""
    N = H.shape[0]
    a = np.asfortranarray(H, dtype=np.complex128).reshape(-1, order="F").copy()
    buf = a.view(np.float64)
    W = np.zeros(N)
    sweeps = lib().orc_zheev(N, _p(buf), _p(W))
    if sweeps < 0:
        raise RuntimeError("oracle zheev did not converge")
    V = buf.view(np.complex128).reshape((N, N), order="F")
    return W, V


def el_diag(prob, k):
    nw = prob.nw
    E = np.zeros(nw)
    U = np.zeros(2 * nw * nw)
    h0 = cplx_view(prob.h0R)
    rc = lib().orc_el_diag(nw, len(prob.Rh), _p(_c(prob.Rh)), _p(_c(prob.degh)), _p(h0),
                           _p(_c(k)), _p(E), _p(U))
    assert rc == 0
    return E, U.view(np.complex128).reshape((nw, nw), order="F")


def ph_diag(prob, q):
    nm = prob.nm
    w = np.zeros(nm)
    ev = np.zeros(2 * nm * nm)
    rc = lib().orc_ph_diag(nm, len(prob.Rf), _p(_c(prob.Rf)), _p(_c(prob.wf)),
                           _p(_c(prob.mat2R)), _p(_c(prob.masses)), _p(_c(q)), _p(w), _p(ev))
    assert rc == 0
    return w, ev.view(np.complex128).reshape((nm, nm), order="F")


def cache_elph(nw, nm, Rp, Re, degE, gW, U1, k1):
    """elPhCached (nRp, nm, nb1, nw) from gW (nRe,nRp,nm,nw,nw) complex."""
    nRp, nRe = len(Rp), len(Re)
    nb1 = U1.shape[1]
    out = np.zeros(2 * nRp * nm * nb1 * nw)
    lib().orc_cache_elph(nw, nm, nRp, nRe, _p(cplx_view(gW)), _p(_c(Re)), _p(_c(degE)),
                         _p(cplx_view(np.asfortranarray(U1).reshape(-1, order="F"))), nb1,
                         _p(_c(k1)), _p(out))
    return out.view(np.complex128).reshape(nRp, nm, nb1, nw)


def coupling_sq(nw, nm, Rp, degP, cached, U2s, e3s, q3s):
    """|g|^2 per q as a list of (nb1, nb2_q, nm) arrays."""
    nRp = len(Rp)
    nb1 = cached.shape[2]
    nq = len(U2s)
    nb2s = np.array([u.shape[1] for u in U2s], dtype=np.int32)
    U2p = np.concatenate([np.asfortranarray(u, dtype=np.complex128).reshape(-1, order="F")
                          for u in U2s])
    e3p = np.concatenate([np.asfortranarray(e, dtype=np.complex128).reshape(-1, order="F")
                          for e in e3s])
    out = np.zeros(int(nb1 * nm * nb2s.sum()))
    lib().orc_coupling_sq(nw, nm, nRp, _p(_c(Rp)), _p(_c(degP)), _p(cplx_view(cached)), nb1, nq,
                          _p(cplx_view(U2p)), nb2s.ctypes.data_as(_ip), _p(cplx_view(e3p)),
                          _p(_c(np.asarray(q3s).reshape(-1))), _p(out))
    res, off = [], 0
    for nb2 in nb2s:
        n = nb1 * nb2 * nm
        res.append(out[off:off + n].reshape((nb1, nb2, nm), order="F"))
        off += n
    return res


def polar_part1(pol, recip, q, e3):
    """polarCorrectionPart1Static at q with eigenvectors e3 (nm, nm); pol: dict of
    positions, born, dielectric, volume, coarse_mesh."""
    nm = e3.shape[0]
    x = np.zeros(2 * nm)
    rc = np.asfortranarray(recip, dtype=np.float64).reshape(-1, order="F").copy()
    cm = np.ascontiguousarray(pol["coarse_mesh"], dtype=np.int32)
    lib().orc_polar_part1(len(pol["positions"]), _p(_c(pol["positions"])), _p(_c(pol["born"])),
                          _p(_c(pol["dielectric"])), _p(rc), C.c_double(pol["volume"]), cm.ctypes.data_as(_ip), nm,
                          _p(_c(q)), _p(cplx_view(np.asfortranarray(e3).reshape(-1, order="F"))), _p(x))
    return x.view(np.complex128)


def coupling_sq_polar(nw, nm, Rp, degP, cached, U2s, e3s, q3s, U1, polar):
    """coupling_sq with the polar term (polar: nq x nm complex, or None)."""
    nRp = len(Rp)
    nb1 = cached.shape[2]
    nq = len(U2s)
    nb2s = np.array([u.shape[1] for u in U2s], dtype=np.int32)
    U2p = np.concatenate([np.asfortranarray(u, dtype=np.complex128).reshape(-1, order="F") for u in U2s])
    e3p = np.concatenate([np.asfortranarray(e, dtype=np.complex128).reshape(-1, order="F") for e in e3s])
    out = np.zeros(int(nb1 * nm * nb2s.sum()))
    pp = None if polar is None else _p(cplx_view(np.asarray(polar).reshape(nq, nm)))
    lib().orc_coupling_sq_polar(nw, nm, nRp, _p(_c(Rp)), _p(_c(degP)), _p(cplx_view(cached)), nb1, nq,
                                _p(cplx_view(U2p)), nb2s.ctypes.data_as(_ip), _p(cplx_view(e3p)),
                                _p(_c(np.asarray(q3s).reshape(-1))),
                                _p(cplx_view(np.asfortranarray(U1).reshape(-1, order="F"))), pp, _p(out))
    res, off = [], 0
    for nb2 in nb2s:
        n = nb1 * nb2 * nm
        res.append(out[off:off + n].reshape((nb1, nb2, nm), order="F"))
        off += n
    return res


def symmetrize_coupling(c, e1, e2, e3):
    a = np.asfortranarray(c, dtype=np.float64).copy()
    nb1, nb2, nm = a.shape
    buf = a.reshape(-1, order="F")
    lib().orc_symmetrize_coupling(_p(buf), nb1, nb2, nm, _p(_c(e1)), _p(_c(e2)), _p(_c(e3)))
    return buf.reshape((nb1, nb2, nm), order="F")


def window(energies, win_min, win_max):
    """(passing count, lo, hi) of one point's bands (lo = hi = -1 when none passes)."""
    e = _c(energies)
    lo, hi = C.c_int(), C.c_int()
    n = lib().orc_window(len(e), _p(e), float(win_min), float(win_max), C.byref(lo), C.byref(hi))
    return n, lo.value, hi.value


def average_point(energies, lw):
    """Degeneracy averaging of one point's linewidths (one row)."""
    out = _c(lw).copy()
    lib().orc_average_point(len(out), _p(_c(energies)), _p(out))
    return out


def gaussian_delta(e, sigma):
    return lib().orc_gaussian_delta(float(e), float(sigma))


def population(is_bose, e, T, mu=0.0):
    return lib().orc_population(int(is_bose), float(e), float(T), float(mu))


def phonon_cutoff():
    return lib().orc_phonon_cutoff()


# ------------------------------------------------------------------ driver
class Oracle:
    """Whole linewidth problem on the CPU (ElScatteringMatrix::builder, case 2)."""

    def __init__(self, prob, gW_flat: np.ndarray):
        L = lib()
        self._keep = []

        def keep(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a

        s = OrcProblem()
        rc = prob.recip
        for j in range(3):
            for i in range(3):
                s.recip[i + 3 * j] = rc[i, j]
        for i in range(3):
            s.mesh[i] = prob.mesh[i]
        s.nw, s.nRh = prob.nw, len(prob.Rh)
        s.Rh, s.degh = _p(keep(prob.Rh)), _p(keep(prob.degh))
        s.h0R = _p(keep(cplx_view(prob.h0R)))
        s.nm, s.nRf = prob.nm, len(prob.Rf)
        s.Rf, s.wf, s.mat2R = _p(keep(prob.Rf)), _p(keep(prob.wf)), _p(keep(prob.mat2R))
        s.masses = _p(keep(prob.masses))
        s.nRe, s.nRp = prob.nRe, prob.nRp
        s.Re, s.degE, s.Rp, s.degP = (_p(keep(prob.Re)), _p(keep(prob.degE)), _p(keep(prob.Rp)),
                                      _p(keep(prob.degP)))
        g = gW_flat if gW_flat.dtype == np.float64 else gW_flat.view(np.float64)
        self._keep.append(g)
        s.gW = _p(g)
        s.nCalc = len(prob.temperatures)
        s.temperatures, s.chemPots = _p(keep(prob.temperatures)), _p(keep(prob.chem_pots))
        s.sigma, s.winMin, s.winMax = prob.sigma, prob.win_min, prob.win_max
        lr = getattr(prob, "long_range", None)
        if lr is not None and lr["nG"] > 0:
            s.natoms, s.nG = int(lr["natoms"]), int(lr["nG"])
            s.atomPos, s.bornCharges = _p(keep(lr["positions"])), _p(keep(lr["born"]))
            s.dielectric, s.gVectors = _p(keep(lr["dielectric"])), _p(keep(lr["gVectors"]))
            s.lrConstant = _p(keep(lr["constant"]))
            s.lrNorm, s.gMax = float(lr["norm"]), float(lr["gMax"])
        s.smearing = int(getattr(prob, "smearing", 0))
        pol = getattr(prob, "polar", None)
        if pol is not None:
            s.polarOn = 1
            s.polarPos, s.polarBorn = _p(keep(pol["positions"])), _p(keep(pol["born"]))
            s.polarEps = _p(keep(pol["dielectric"]))
            s.polarVolume = float(pol["volume"])
            for i in range(3):
                s.polarCoarse[i] = int(pol["coarse_mesh"][i])
        self._s = s
        self.h = L.orc_setup(C.byref(s))
        if not self.h:
            raise RuntimeError("oracle setup failed")
        self.prob = prob
        self.num_active = L.orc_num_active(self.h)
        self.num_states = L.orc_num_states(self.h)
        self.active = np.zeros(self.num_active, dtype=np.int32)
        self.band_lo = np.zeros(self.num_active, dtype=np.int32)
        self.band_hi = np.zeros(self.num_active, dtype=np.int32)
        self.energies = np.zeros(self.num_states)
        L.orc_get_window(self.h, self.active.ctypes.data_as(_ip), self.band_lo.ctypes.data_as(_ip),
                         self.band_hi.ctypes.data_as(_ip), _p(self.energies))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def phel_linewidths(self, q_mesh, average=True, k1_list=None):
        """Phonon linewidths (nCalc, nq, nm) of PhElScatteringMatrix::builder
        on the q mesh, from this problem's electrons, phonons and coupling;
        k1_list: outer points (active indices), all when None.  The outer
        loop's wall time is left in self.phel_seconds."""
        nq = int(np.prod(q_mesh))
        out = np.zeros(len(self.prob.temperatures) * nq * self.prob.nm)
        qm = np.ascontiguousarray(q_mesh, dtype=np.int32)
        sec = np.zeros(1)
        if k1_list is None:
            n, kl = -1, None
        else:
            k = np.ascontiguousarray(k1_list, dtype=np.int32)
            n, kl = len(k), k.ctypes.data_as(_ip)
        pairs = lib().orc_phel(C.byref(self._s), qm.ctypes.data_as(_ip), n, kl, int(bool(average)), _p(out),
                               _p(sec))
        if pairs < 0:
            raise RuntimeError("oracle ph-el failed")
        self.phel_pairs = pairs
        self.phel_seconds = float(sec[0])
        return out.reshape(len(self.prob.temperatures), nq, self.prob.nm)

    def eigvecs(self, a):
        nb = int(self.band_hi[a] - self.band_lo[a] + 1)
        out = np.zeros(2 * self.prob.nw * nb)
        lib().orc_get_eigvecs(self.h, int(a), _p(out))
        return out.view(np.complex128).reshape((self.prob.nw, nb), order="F")

    def phonons(self):
        nk = int(np.prod(self.prob.mesh))
        nm = self.prob.nm
        w = np.zeros(nk * nm)
        ev = np.zeros(2 * nk * nm * nm)
        ins = C.create_string_buffer(nk)
        lib().orc_get_phonons(self.h, _p(w), _p(ev), ins)
        return (w.reshape(nk, nm), ev.view(np.complex128).reshape(nk, nm, nm).transpose(0, 2, 1),
                np.frombuffer(ins.raw, dtype=np.uint8).astype(bool))

    def linewidths(self, k1_list=None, k2_range=None, average=True):
        """Linewidths (nCalc, numStates) in Ry for the listed outer points."""
        nc = len(self.prob.temperatures)
        lw = np.zeros(nc * self.num_states)
        if k1_list is None:
            nk1, ptr = -1, None
        else:
            arr = np.ascontiguousarray(k1_list, dtype=np.int32)
            self._k1 = arr
            nk1, ptr = len(arr), arr.ctypes.data_as(_ip)
        b, e = (0, -1) if k2_range is None else k2_range
        pairs = lib().orc_linewidth_pairs(self.h, nk1, ptr, int(b), int(e), _p(lw))
        if pairs < 0:
            raise RuntimeError("oracle linewidth failed")
        if average:
            lib().orc_degeneracy_average(self.h, _p(lw))
        self.last_pairs = pairs
        return lw.reshape(nc, self.num_states)


def adaptive_delta(energy, v, recip, mesh):
    """AdaptiveGaussianDeltaFunction::getSmearing for v = v1 - v2."""
    rc = np.asfortranarray(recip, dtype=np.float64).reshape(-1, order="F").copy()
    m = np.ascontiguousarray(mesh, dtype=np.int32)
    return lib().orc_adaptive_delta(float(energy), _p(_c(v)), _p(rc), m.ctypes.data_as(_ip))


def el_group_velocities(prob, k):
    """(nw, 3) group velocities of the single-point velocity operator at k."""
    out = np.zeros(3 * prob.nw)
    rc = lib().orc_el_group_velocities(prob.nw, len(prob.Rh), _p(_c(prob.Rh)), _p(_c(prob.degh)),
                                       _p(cplx_view(prob.h0R)), _p(_c(k)), _p(out))
    assert rc == 0
    return out.reshape(prob.nw, 3)


def accounted_bytes(nw, nm, nRp, nRe, nb1_cached=0):
    """InteractionElPhWan::getDeviceMemoryUsage (interaction_elph.cpp:703-707):
    16 bytes per element of the coupling tensor and of elPhCached, 8 per
    element of the R vectors and degeneracies."""
    return 16.0 * (nRe * nRp * nm * nw * nw + nRp * nm * nb1_cached * nw) + 8.0 * (nRp + 3 * nRp + 3 * nRe + nRe)


def estimate_num_batches(nw, nm, nRp, nk2, nb1, available):
    """InteractionElPhWan::estimateNumBatches (interaction_elph.cpp:458-490):
    ceil(nk2 * per-q bytes / available); None for the "Insufficient memory!"
    error (one q does not fit)."""
    import math
    evs = 16 * (nw * nw + nm * nm)
    phase = 16 * nRp
    g3 = 2 * 16 * nm * nb1 * nw
    g4 = 2 * 16 * nm * nb1 * nw
    gfinal = 2 * 16 * nm * nb1 * nw
    coupling = 16 * nb1 * nw * nm
    polar = 16 * nm * nb1 * nw
    usage = nk2 * (evs + polar + max(phase + g3, g3 + g4, g4 + gfinal, gfinal + coupling))
    if nk2 > 0 and available < usage / nk2:
        return None
    return int(math.ceil(usage / available))
