"""phoebe_amd - MI355X (gfx950) electron-phonon linewidth path.

Python front-end over the C ABI of libphx_elph.so (include/phx_elph.h).  The
classes mirror the reference interface of the hot path:

* InteractionElPhWan  (src/interaction/interaction_elph.h:30-219):
  cacheElPh / calcCouplingSquared / getCouplingSquared /
  getCouplingDimensions / estimateNumBatches;
* zheev_batched       (kokkosZHEEV, src/common_kokkos.h:49);
* el_bloch_diag / ph_bloch_diag (kokkosBatchedDiagonalizeFromCoordinates,
  src/harmonic/harmonic.h:87-91);
* ElectronLinewidths  (ElScatteringMatrix::builder, switchCase 2,
  src/bte/el_scattering.cpp:38-473);
* ElPhFile / write_elph_file  (Phoebe's prefix.phoebe.elph.hdf5,
  interaction_elph_parsing.cpp:161-670, qe_to_phoebe_utils.cpp:377-720),
  read and written by the library's own HDF5 subset (no libhdf5, no GPU).

The HIP library is required: importing works without it (so that build
machines can import the package), but every compute call raises if the
library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PHX_LIBRARY: load the library from another path (a build of different options)
LIB_PATH = os.environ.get("PHX_LIBRARY") or os.path.join(_HERE, "libphx_elph.so")

PHX_FP64, PHX_FP32, PHX_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PHX_FP64, "fp32": PHX_FP32, "bf16x3": PHX_BF16X3}
TRANSFORM_DIRECT, TRANSFORM_SEPARABLE = 0, 1
TRANSFORMS = {"direct": TRANSFORM_DIRECT, "separable": TRANSFORM_SEPARABLE}
GAMMA_BATCHED, GAMMA_POINT = 0, 1
SMEAR_GAUSSIAN, SMEAR_ADAPTIVE = 0, 1

# Named configurations of the linewidth builder: (precision, transform,
# energy screen in units of sigma (0: off), line batching: least outer points
# of a mesh line for the line form of cacheElPh (0: off)).  "fp64" is the
# plain form (the reference algorithm in FP64); CHOSEN is the fastest form
# whose linewidths all stay within 1e-6 relative of FP64.
FORMS = {
    "fp64": ("fp64", "direct", 0.0, 0),
    "fp32": ("fp32", "direct", 0.0, 0),
    "sep64": ("fp64", "separable", 0.0, 0),
    "sep32": ("fp32", "separable", 0.0, 0),
    "sep32s": ("fp32", "separable", 7.0, 0),
    "sep64s": ("fp64", "separable", 7.0, 0),
    "line32s": ("fp32", "separable", 7.0, 8),
    "line64s": ("fp64", "separable", 7.0, 8),
    "line64s6": ("fp64", "separable", 6.0, 8),
    "line32s6": ("fp32", "separable", 6.0, 8),
    "line3s6": ("bf16x3", "separable", 6.0, 8),
}
CHOSEN = "line32s6"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


class PhxError(RuntimeError):
    pass


class PhxProblem(C.Structure):
    _fields_ = [
        ("recip", C.c_double * 9), ("mesh", C.c_int * 3),
        ("nw", C.c_int), ("nRh", C.c_int), ("Rh", _dp), ("degh", _dp), ("h0R", _dp),
        ("nm", C.c_int), ("nRf", C.c_int), ("Rf", _dp), ("wf", _dp), ("mat2R", _dp),
        ("masses", _dp),
        ("nRe", C.c_int), ("nRp", C.c_int), ("Re", _dp), ("degE", _dp), ("Rp", _dp),
        ("degP", _dp), ("gW", _dp), ("gwDecay", _dp), ("gwSeed", C.c_ulonglong),
        ("nCalc", C.c_int), ("temperatures", _dp), ("chemPots", _dp),
        ("sigma", C.c_double), ("winMin", C.c_double), ("winMax", C.c_double),
        ("natoms", C.c_int), ("nG", C.c_int), ("atomPos", _dp), ("bornCharges", _dp),
        ("dielectric", _dp), ("gVectors", _dp), ("lrConstant", _dp),
        ("lrNorm", C.c_double), ("gMax", C.c_double),
        ("smearing", C.c_int),
    ]


EXPORTS = {
    "phx_last_error": (C.c_char_p, []),
    "phx_version": (C.c_char_p, []),
    "phx_device_count": (C.c_int, [_ip]),
    "phx_bloch_diag_timed": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                       C.c_int, _dp, _dp, _dp]),
    "phx_zheev_batched": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _ip]),
    "phx_zheev_batched_dev": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "phx_el_bloch_diag": (C.c_int, [C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "phx_el_bloch_diag_noalloc": (C.c_int, [C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "phx_ph_bloch_diag": (C.c_int, [C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int,
                                    _dp, _dp]),
    "phx_elph_create": (C.c_int, [C.POINTER(C.c_void_p), _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _dp, _dp, _dp, _dp, C.c_int]),
    "phx_elph_create_synthetic": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int,
                                            C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_ulonglong,
                                            C.c_int]),
    "phx_elph_cache": (C.c_int, [C.c_void_p, _dp, C.c_int, _dp]),
    "phx_elph_coupling_sq": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip, _dp, _dp, _dp, _dp]),
    "phx_elph_get_coupling_dims": (C.c_int, [C.c_void_p, _ip]),
    "phx_elph_estimate_num_batches": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
    "phx_elph_device_memory_usage": (C.c_int, [C.c_void_p, _dp]),
    "phx_polar_in_use": (C.c_int, [_dp, C.c_int]),
    "phx_elph_set_polar": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phx_elph_polar_part1": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "phx_lw_set_polar": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phx_elph_destroy": (C.c_int, [C.c_void_p]),
    "phx_lw_setup": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PhxProblem), C.c_int, C.c_int,
                               C.c_int]),
    "phx_lw_num_active": (C.c_int, [C.c_void_p]),
    "phx_lw_num_states": (C.c_int, [C.c_void_p]),
    "phx_lw_get_window": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _dp]),
    "phx_lw_get_phonons": (C.c_int, [C.c_void_p, _dp, C.c_char_p]),
    "phx_lw_run": (C.c_longlong, [C.c_void_p, C.c_int, _ip, C.c_int, C.c_int]),
    "phx_lw_get_linewidths": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "phx_lw_reset": (C.c_int, [C.c_void_p]),
    "phx_lw_sync": (C.c_int, [C.c_void_p]),
    "phx_lw_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_set_screen": (C.c_int, [C.c_void_p, C.c_double]),
    "phx_lw_set_line_batching": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_set_panel_chunk": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_screen_collect": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_ulonglong]),
    "phx_lw_screen_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_ulonglong,
                                      C.c_void_p]),
    "phx_lw_set_learned_screen": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_ulonglong]),
    "phx_lw_screen_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phx_lw_screen_reset_stats": (C.c_int, [C.c_void_p]),
    "phx_lw_screen_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "phx_mlp_num_params": (C.c_int, []),
    "phx_lw_stage_times": (C.c_int, [C.c_void_p, C.c_void_p]),
    "phx_lw_plan": (C.c_longlong, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int]),
    "phx_lw_plan_run": (C.c_longlong, [C.c_void_p, C.c_int]),
    "phx_lw_set_batching": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "phx_lw_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "phx_stream_copy": (C.c_int, [C.c_ulonglong, C.c_int, C.c_void_p]),
    "phx_mlp_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "phx_lw_kernel_times": (C.c_int, [C.c_void_p, _dp]),
    "phx_lw_stream": (C.c_void_p, [C.c_void_p]),
    "phx_lw_copy_gw": (C.c_int, [C.c_void_p, _dp]),
    "phx_lw_destroy": (C.c_int, [C.c_void_p]),
    "phx_synth_gw_host": (C.c_int, [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_ulonglong]),
    "phx_elph_file_open": (C.c_int, [C.POINTER(C.c_void_p), C.c_char_p]),
    "phx_elph_file_header": (C.c_int, [C.c_void_p, _ip]),
    "phx_elph_file_bravais": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "phx_elph_file_read_gw": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "phx_elph_file_close": (C.c_int, [C.c_void_p]),
    "phx_elph_file_write": (C.c_int, [C.c_char_p, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp]),
    "phx_elph_create_from_file": (C.c_int, [C.POINTER(C.c_void_p), C.c_char_p, C.c_int]),
    "phx_lw_setup_file": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PhxProblem), C.c_char_p,
                                    C.c_int, C.c_int, C.c_int]),
    "phx_ph_model_read_qe": (C.c_int, [C.POINTER(C.c_void_p), C.c_char_p, C.c_char_p]),
    "phx_ph_model_dims": (C.c_int, [C.c_void_p, _ip]),
    "phx_ph_model_crystal": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "phx_ph_model_force_constants": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "phx_ph_model_long_range": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "phx_ph_model_diag": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int, _dp, _dp]),
    "phx_ph_model_destroy": (C.c_int, [C.c_void_p]),
    "phx_el_model_read_w90": (C.c_int, [C.POINTER(C.c_void_p), C.c_char_p, _dp]),
    "phx_el_model_dims": (C.c_int, [C.c_void_p, _ip]),
    "phx_el_model_arrays": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "phx_el_model_diag": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "phx_el_model_destroy": (C.c_int, [C.c_void_p]),
    "phx_problem_set_models": (C.c_int, [C.POINTER(PhxProblem), C.c_void_p, C.c_void_p]),
    "phx_ph_model_diag_velocities": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int, _dp, _dp, _dp]),
    "phx_phel_setup": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PhxProblem), C.c_void_p, _ip]),
    "phx_phel_num_active": (C.c_int, [C.c_void_p]),
    "phx_phel_run": (C.c_longlong, [C.c_void_p, C.c_int, _ip]),
    "phx_phel_get_linewidths": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "phx_phel_get_phonons": (C.c_int, [C.c_void_p, _dp]),
    "phx_phel_reset": (C.c_int, [C.c_void_p]),
    "phx_phel_destroy": (C.c_int, [C.c_void_p]),
    "phx_phel_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_set_split_last_stage": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_set_epilogue_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_set_fused_y_stage": (C.c_int, [C.c_void_p, C.c_int]),
    "phx_lw_last_stage_fused_y": (C.c_int, [C.c_void_p, _ip]),
    "phx_lw_last_stage_split_times": (C.c_int, [C.c_void_p, _dp]),
    "phx_lw_last_stage_split_bytes": (C.c_int, [C.c_void_p, _dp]),
    "phx_phel_kernel_times": (C.c_int, [C.c_void_p, _dp]),
    "phx_el_model_diag_velocities": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int, _dp, _dp, _dp]),
}


def lib():
    """The loaded HIP library; raises PhxError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PhxError(f"{LIB_PATH} is missing: run `python -m phoebe_amd.build` first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise PhxError(lib().phx_last_error().decode())
    return rc


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _cview(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


def _fortran_flat(m):
    return np.asfortranarray(m, dtype=np.complex128).reshape(-1, order="F")


# ------------------------------------------------------------------ eigensolver / Bloch
def device_count() -> int:
    """HIP devices visible to this process (0 without a GPU)."""
    n = C.c_int(0)
    _check(lib().phx_device_count(C.byref(n)))
    return n.value


def zheev_batched(A: np.ndarray):
    """Eigen-decomposition of a stack of Hermitian matrices A[M, N, N].

    Returns (W[M, N] ascending, V[M, N, N]) with V[i][:, j] the eigenvectors.
    """
    A = np.asarray(A, dtype=np.complex128)
    M, N, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1))).view(np.float64).copy()
    W = np.zeros((M, N))
    info = np.zeros(M, dtype=np.int32)
    _check(lib().phx_zheev_batched(M, N, _p(buf), _p(W), info.ctypes.data_as(_ip)))
    V = buf.view(np.complex128).reshape(M, N, N).transpose(0, 2, 1)
    return W, V


def el_bloch_diag(kc, R, deg, h0R, noalloc=False):
    """Energies (nk, nw) and eigenvectors (nk, nw, nw) of H(k).  noalloc: the
    entry point that works in the library's fixed device scratch."""
    kc = _f64(kc).reshape(-1, 3)
    nk, nw = len(kc), h0R.shape[1]
    E = np.zeros((nk, nw))
    U = np.zeros(2 * nk * nw * nw)
    fn = lib().phx_el_bloch_diag_noalloc if noalloc else lib().phx_el_bloch_diag
    _check(fn(nk, _p(kc), nw, len(R), _p(_f64(R)), _p(_f64(deg)), _p(_cview(h0R)), _p(E), _p(U)))
    return E, U.view(np.complex128).reshape(nk, nw, nw).transpose(0, 2, 1)


def ph_bloch_diag(qc, R, w, mat2R, masses, mass_scaling=True):
    qc = _f64(qc).reshape(-1, 3)
    nq, nm = len(qc), mat2R.shape[1]
    om = np.zeros((nq, nm))
    ev = np.zeros(2 * nq * nm * nm)
    _check(lib().phx_ph_bloch_diag(nq, _p(qc), nm, len(R), _p(_f64(R)), _p(_f64(w)),
                                   _p(_f64(mat2R)), _p(_f64(masses)), int(bool(mass_scaling)),
                                   _p(om), _p(ev)))
    return om, ev.view(np.complex128).reshape(nq, nm, nm).transpose(0, 2, 1)


def bloch_diag_timed(kind, kc, R, weight, coef, masses=None, reps=3):
    """Bloch build + batched eigensolve timed on device-resident inputs:
    returns (W, V, ms_bloch, ms_eig) with ms per launch."""
    kc = _f64(kc).reshape(-1, 3)
    nk, n = len(kc), coef.shape[1]
    W = np.zeros((nk, n))
    U = np.zeros(2 * nk * n * n)
    ms = np.zeros(2)
    ph = kind == "ph"
    cf = _f64(coef) if ph else _cview(coef)
    _check(lib().phx_bloch_diag_timed(int(ph), nk, _p(kc), n, len(R), _p(_f64(R)), _p(_f64(weight)), _p(cf),
                                      _p(_f64(masses)) if ph else None, int(reps), _p(ms), _p(W), _p(U)))
    return W, U.view(np.complex128).reshape(nk, n, n).transpose(0, 2, 1), float(ms[0]), float(ms[1])


def stream_copy_ms(nbytes, reps=5):
    """Average milliseconds of a plain device copy of nbytes (HBM reference)."""
    out = np.zeros(1)
    _check(lib().phx_stream_copy(int(nbytes), int(reps), _p(out)))
    return float(out[0])


def mlp_gradients(params, X, t):
    """(loss, grads, y) of the screen's MLP on the GPU forward / backward kernels."""
    n = len(t)
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32)
    tc = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(lib().phx_mlp_num_params(), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    y = np.zeros(n, dtype=np.float32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    _check(lib().phx_mlp_gradients(vp(P), vp(Xc), vp(tc), n, vp(g), vp(loss), vp(y)))
    return float(loss[0]), g, y


# ------------------------------------------------------------------ el-ph file
HEADER_KEYS = ("fileFormat", "numElectrons", "numSpin", "numElBands", "numPhModes", "nRe", "nRp")


def _path(path):
    return os.fsencode(os.fspath(path))


class ElPhFile:
    """Phoebe's el-ph coupling file, prefix.phoebe.elph.hdf5, formats 1 and 2.

    Mirrors parseHDF5 / parseHeaderHDF5 / parseHDF5V1 / parseHDF5V2
    (interaction_elph_parsing.cpp:161-670).  Host code only.
    """

    def __init__(self, path):
        self.path = os.fspath(path)
        self._h = C.c_void_p()
        _check(lib().phx_elph_file_open(C.byref(self._h), _path(path)))
        v = np.zeros(13, dtype=np.int32)
        _check(lib().phx_elph_file_header(self._h, v.ctypes.data_as(_ip)))
        self.header = {k: int(x) for k, x in zip(HEADER_KEYS, v[:7])}
        self.header["kMesh"] = tuple(int(x) for x in v[7:10])
        self.header["qMesh"] = tuple(int(x) for x in v[10:13])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().phx_elph_file_close(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def bravais(self):
        """(elR (nRe, 3), elDeg (nRe,), phR (nRp, 3), phDeg (nRp,)), Bohr."""
        h = self.header
        eR, eD = np.zeros((h["nRe"], 3)), np.zeros(h["nRe"])
        pR, pD = np.zeros((h["nRp"], 3)), np.zeros(h["nRp"])
        _check(lib().phx_elph_file_bravais(self._h, _p(eR), _p(eD), _p(pR), _p(pD)))
        return eR, eD, pR, pD

    def read_gw(self, re_begin=0, re_count=None):
        """gW of R_e in [re_begin, re_begin + re_count) as complex
        (re_count, nRp, nm, nw, nw) (the device layout of phx_elph_create)."""
        h = self.header
        if re_count is None:
            re_count = h["nRe"] - re_begin
        out = np.zeros((re_count, h["nRp"], h["numPhModes"], h["numElBands"], h["numElBands"]),
                       dtype=np.complex128)
        _check(lib().phx_elph_file_read_gw(self._h, int(re_begin), int(re_count),
                                           _p(out.view(np.float64))))
        return out


def write_elph_file(path, gW, elR, elDeg, phR, phDeg, numElectrons, numSpin=1,
                    kMesh=(0, 0, 0), qMesh=(0, 0, 0), fileFormat=1):
    """writeElPhCouplingHDF5v1 / v2 (qe_to_phoebe_utils.cpp:446-720): gW complex
    (nRe, nRp, nm, nw, nw); elR (nRe, 3), phR (nRp, 3)."""
    g = np.ascontiguousarray(gW, dtype=np.complex128)
    nRe, nRp, nm, nw, nw2 = g.shape
    eR, eD, pR, pD = (_f64(a) for a in (elR, elDeg, phR, phDeg))
    if nw2 != nw or eR.shape != (nRe, 3) or pR.shape != (nRp, 3) or eD.size != nRe or pD.size != nRp:
        raise PhxError("write_elph_file: inconsistent shapes")
    v = np.array([numElectrons, numSpin, nw, nm, nRe, nRp, *kMesh, *qMesh], dtype=np.int32)
    _check(lib().phx_elph_file_write(_path(path), int(fileFormat), v.ctypes.data_as(_ip), _p(eR),
                                     _p(eD), _p(pR), _p(pD), _p(g.view(np.float64))))


# ------------------------------------------------------------------ polar term
class PhxPolar(C.Structure):
    _fields_ = [("natoms", C.c_int), ("atomPos", _dp), ("bornCharges", _dp), ("dielectric", _dp),
                ("recip", C.c_double * 9), ("volume", C.c_double), ("coarseMesh", C.c_int * 3)]


def polar_struct(positions, born, dielectric, recip, volume, coarse_mesh):
    """phx_polar from atomic positions (natoms, 3) Bohr, Born charges
    Z(atom, i, pol) (natoms, 3, 3), the dielectric tensor (3, 3), the
    reciprocal cell (columns b_i), the cell volume (Bohr^3) and the force
    constants' q grid.  Returns (struct, arrays it points into)."""
    pos, z, eps = _f64(positions), _f64(born), _f64(dielectric)
    if pos.ndim != 2 or pos.shape[1] != 3 or z.shape != (len(pos), 3, 3) or eps.shape != (3, 3):
        raise PhxError("polar_struct: inconsistent shapes")
    st = PhxPolar()
    st.natoms = len(pos)
    st.atomPos, st.bornCharges, st.dielectric = _p(pos), _p(z), _p(eps)
    rc = np.asarray(recip, dtype=np.float64)
    for j in range(3):
        for i in range(3):
            st.recip[i + 3 * j] = rc[i, j]
    st.volume = float(volume)
    for i in range(3):
        st.coarseMesh[i] = int(coarse_mesh[i])
    return st, (pos, z, eps)


def polar_in_use(dielectric, num_species):
    """The reference's usePolarCorrection test (interaction_elph.cpp:24-32)."""
    return bool(lib().phx_polar_in_use(_p(_f64(dielectric)), int(num_species)))


# ------------------------------------------------------------------ InteractionElPhWan
class InteractionElPhWan:
    """Wannier electron-phonon coupling on the GPU (interaction_elph.h:30-219)."""

    def __init__(self, couplingWannier=None, elBravaisVectors=None, elDegeneracies=None,
                 phBravaisVectors=None, phDegeneracies=None, precision="fp64", *,
                 synthetic=None):
        """couplingWannier: complex (nRe, nRp, nm, nw, nw) (the reference's
        device view of its Eigen tensor); or synthetic=(nw, nm, decay, seed)."""
        self._h = C.c_void_p()
        if isinstance(couplingWannier, (str, bytes, os.PathLike)):
            # InteractionElPhWan::parse (interaction_elph_parsing.cpp:672-692)
            prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
            _check(lib().phx_elph_create_from_file(C.byref(self._h), _path(couplingWannier), prec))
            dims = self.getCouplingDimensions()
            self.numElBands, self.numPhBands = int(dims[3]), int(dims[2])
            self._cache = []
            self._nb1 = None
            return
        elR, phR = _f64(elBravaisVectors), _f64(phBravaisVectors)
        elD, phD = _f64(elDegeneracies), _f64(phDegeneracies)
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        if synthetic is not None:
            nw, nm, decay, seed = synthetic
            _check(lib().phx_elph_create_synthetic(C.byref(self._h), nw, nm, len(phR), len(elR),
                                                   _p(elR), _p(elD), _p(phR), _p(phD),
                                                   _p(_f64(decay)), int(seed), prec))
        else:
            g = np.ascontiguousarray(couplingWannier, dtype=np.complex128)
            nRe, nRp, nm, nw, _ = g.shape
            _check(lib().phx_elph_create(C.byref(self._h), _p(g.view(np.float64)), nw, nm, nRp, nRe,
                                         _p(elR), _p(elD), _p(phR), _p(phD), prec))
            self.numElBands, self.numPhBands = nw, nm
        dims = self.getCouplingDimensions()
        self.numElBands, self.numPhBands = int(dims[3]), int(dims[2])
        self._cache = []
        self._nb1 = None

    @classmethod
    def parse(cls, path, precision="fp64"):
        """The interaction read from Phoebe's el-ph file, gW streamed into HBM."""
        return cls(os.fspath(path), precision=precision)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().phx_elph_destroy(self._h)
            self._h = C.c_void_p()

    def cacheElPh(self, eigvec1, k1C):
        U1 = _fortran_flat(eigvec1)
        self._nb1 = int(eigvec1.shape[1])
        _check(lib().phx_elph_cache(self._h, _p(U1.view(np.float64)), self._nb1, _p(_f64(k1C))))

    usePolarCorrection = False

    def setPolar(self, positions, born, dielectric, recip, volume, coarse_mesh, num_species):
        """The Crystal / PhononH0 fields of the polar term (interaction_elph.cpp:
        24-32, 140-151); the term is used when the reference would use it."""
        st, self._polar_keep = polar_struct(positions, born, dielectric, recip, volume, coarse_mesh)
        _check(lib().phx_elph_set_polar(self._h, C.byref(st)))
        self.usePolarCorrection = polar_in_use(dielectric, num_species)

    def polarCorrectionPart1(self, q3, ev3):
        """x(nu) of the Froehlich term at q3 (interaction_elph.cpp:138-203)."""
        return self.polarCorrectionPart1Batch([q3], [ev3])[0]

    def polarCorrectionPart1Batch(self, q3s, ev3s):
        nq = len(q3s)
        nm = self.numPhBands
        q = _f64(np.asarray(q3s).reshape(nq, 3))
        e3 = np.concatenate([_fortran_flat(e) for e in ev3s])
        out = np.zeros(2 * nq * nm)
        _check(lib().phx_elph_polar_part1(self._h, nq, _p(q), _p(e3.view(np.float64)), _p(out)))
        return out.view(np.complex128).reshape(nq, nm)

    def calcCouplingSquared(self, eigvec1, eigvecs2, eigvecs3, q3Cs, polarData=None):
        """polarData: per q the polarCorrectionPart1 vector (nm), used when
        usePolarCorrection (interaction_elph.cpp:285-312, 406-413)."""
        nb1 = int(eigvec1.shape[1])
        if nb1 != self._nb1:
            raise PhxError("eigvec1 does not match the cached k1")
        nq = len(eigvecs2)
        nb2s = np.array([u.shape[1] for u in eigvecs2], dtype=np.int32)
        nm = self.numPhBands
        if nq == 0:
            self._cache = []
            return
        U2 = np.concatenate([_fortran_flat(u) for u in eigvecs2])
        e3 = np.concatenate([_fortran_flat(e) for e in eigvecs3])
        q3 = _f64(np.asarray(q3Cs).reshape(nq, 3))
        out = np.zeros(int(nb1 * nm * nb2s.sum()))
        pol = None
        if self.usePolarCorrection:
            if polarData is None or len(polarData) != nq:
                raise PhxError("calcCouplingSquared: polarData needed for every q3")
            pol = np.ascontiguousarray(np.asarray(polarData, dtype=np.complex128).reshape(nq, nm)).view(np.float64)
        _check(lib().phx_elph_coupling_sq(self._h, nq, _p(U2.view(np.float64)),
                                          nb2s.ctypes.data_as(_ip), _p(e3.view(np.float64)),
                                          _p(q3), None if pol is None else _p(pol), _p(out)))
        self._cache, off = [], 0
        for nb2 in nb2s:
            n = nb1 * nb2 * nm
            self._cache.append(out[off:off + n].reshape((nb1, nb2, nm), order="F"))
            off += n

    def getCouplingSquared(self, ik2):
        return self._cache[ik2]

    def getCouplingDimensions(self):
        d = np.zeros(5, dtype=np.int32)
        _check(lib().phx_elph_get_coupling_dims(self._h, d.ctypes.data_as(_ip)))
        return d

    def estimateNumBatches(self, nk2, nb1, available_bytes=0.0):
        """Batches of inner points (interaction_elph.cpp:458-490); available_bytes:
        the caller's available device memory, or <= 0 for MAXMEM (env, GB,
        default 16) minus this object's accounted usage."""
        return _check(lib().phx_elph_estimate_num_batches(self._h, int(nk2), int(nb1),
                                                          float(available_bytes)))

    def getDeviceMemoryUsage(self):
        """Bytes the reference accounts for this object (interaction_elph.cpp:703-707)."""
        out = C.c_double(0.0)
        _check(lib().phx_elph_device_memory_usage(self._h, C.byref(out)))
        return out.value


# ------------------------------------------------------------------ linewidth builder
def make_problem_struct(prob, gW=None):
    """Fill a PhxProblem from a phoebe_amd.synth.Problem; keeps arrays alive."""
    keep = []

    def k(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _p(a)

    s = PhxProblem()
    rc = prob.recip
    for j in range(3):
        for i in range(3):
            s.recip[i + 3 * j] = rc[i, j]
    for i in range(3):
        s.mesh[i] = prob.mesh[i]
    s.nw, s.nRh = prob.nw, len(prob.Rh)
    s.Rh, s.degh, s.h0R = k(prob.Rh), k(prob.degh), k(_cview(prob.h0R))
    s.nm, s.nRf = prob.nm, len(prob.Rf)
    s.Rf, s.wf, s.mat2R, s.masses = k(prob.Rf), k(prob.wf), k(prob.mat2R), k(prob.masses)
    s.nRe, s.nRp = prob.nRe, prob.nRp
    s.Re, s.degE, s.Rp, s.degP = k(prob.Re), k(prob.degE), k(prob.Rp), k(prob.degP)
    if gW is not None:
        g = gW.view(np.float64) if gW.dtype == np.complex128 else gW
        keep.append(g)
        s.gW = _p(g)
    if getattr(prob, "gw_decay", None) is not None:
        s.gwDecay = k(prob.gw_decay)
    s.gwSeed = int(getattr(prob, "gw_seed", 0))
    s.smearing = int(getattr(prob, "smearing", 0))
    lr = getattr(prob, "long_range", None)
    if lr is not None and lr["nG"] > 0:
        s.natoms, s.nG = int(lr["natoms"]), int(lr["nG"])
        s.atomPos, s.bornCharges = k(lr["positions"]), k(lr["born"])
        s.dielectric, s.gVectors, s.lrConstant = k(lr["dielectric"]), k(lr["gVectors"]), k(lr["constant"])
        s.lrNorm, s.gMax = float(lr["norm"]), float(lr["gMax"])
    s.nCalc = len(prob.temperatures)
    s.temperatures, s.chemPots = k(prob.temperatures), k(prob.chem_pots)
    s.sigma, s.winMin, s.winMax = prob.sigma, prob.win_min, prob.win_max
    return s, keep


class ElectronLinewidths:
    """Electron linewidths from el-ph scattering on the GPU.

    Mirrors ElScatteringMatrix::builder in the linewidth case
    (el_scattering.cpp:38-473): band structure and energy window, phonons on
    the q-set, cacheElPh + calcCouplingSquared + symmetrizeCoupling +
    Gaussian-delta reduction per outer point, degeneracy averaging.
    """

    def __init__(self, prob, gW=None, precision="fp64", k1_batch=32, transform="direct",
                 elph_file=None):
        """gW: host coupling (complex (nRe, nRp, nm, nw, nw)); None with elph_file
        unset uses the problem's synthetic recipe; elph_file: path of Phoebe's
        el-ph file, whose R vectors and gW replace the problem's."""
        self.prob = prob
        self._s, self._keep = make_problem_struct(prob, gW)
        self._h = C.c_void_p()
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        tr = TRANSFORMS[transform] if isinstance(transform, str) else int(transform)
        self.precision = {v: k for k, v in PRECISIONS.items()}[prec]
        self.transform = {v: k for k, v in TRANSFORMS.items()}[tr]
        if elph_file is not None:
            _check(lib().phx_lw_setup_file(C.byref(self._h), C.byref(self._s), _path(elph_file),
                                           prec, int(k1_batch), tr))
        else:
            _check(lib().phx_lw_setup(C.byref(self._h), C.byref(self._s), prec, int(k1_batch), tr))
        self.num_active = lib().phx_lw_num_active(self._h)
        self.num_states = lib().phx_lw_num_states(self._h)
        self.active = np.zeros(self.num_active, dtype=np.int32)
        self.band_lo = np.zeros(self.num_active, dtype=np.int32)
        self.band_hi = np.zeros(self.num_active, dtype=np.int32)
        self.energies = np.zeros(self.num_states)
        _check(lib().phx_lw_get_window(self._h, self.active.ctypes.data_as(_ip),
                                       self.band_lo.ctypes.data_as(_ip),
                                       self.band_hi.ctypes.data_as(_ip), _p(self.energies)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().phx_lw_destroy(self._h)
            self._h = C.c_void_p()

    def phonons(self):
        nk = int(np.prod(self.prob.mesh))
        om = np.zeros(nk * self.prob.nm)
        mask = C.create_string_buffer(nk)
        _check(lib().phx_lw_get_phonons(self._h, _p(om), mask))
        return om.reshape(nk, self.prob.nm), np.frombuffer(mask.raw, dtype=np.uint8).astype(bool)

    def run(self, k1_list=None, k2_range=None):
        """Enqueue the outer points k1_list (active indices; None = all)."""
        if k1_list is None:
            nk1, ptr = -1, None
        else:
            self._k1 = np.ascontiguousarray(k1_list, dtype=np.int32)
            nk1, ptr = len(self._k1), self._k1.ctypes.data_as(_ip)
        b, e = (0, -1) if k2_range is None else k2_range
        pairs = lib().phx_lw_run(self._h, nk1, ptr, int(b), int(e))
        if pairs < 0:
            raise PhxError(lib().phx_last_error().decode())
        return pairs

    def plan(self, k1_list=None, k2_range=None):
        """Prepare a whole sweep (device metadata per batch); returns the batch count."""
        k1 = None if k1_list is None else np.ascontiguousarray(k1_list, dtype=np.int32)
        b, e = (0, -1) if k2_range is None else k2_range
        n = lib().phx_lw_plan(self._h, 0 if k1 is None else len(k1),
                              None if k1 is None else k1.ctypes.data_as(_ip), int(b), int(e))
        if n < 0:
            raise PhxError(lib().phx_last_error().decode())
        return n

    def plan_run(self, graph=True):
        """Enqueue the planned sweep (as one captured hipGraph when graph); asynchronous."""
        pairs = lib().phx_lw_plan_run(self._h, int(bool(graph)))
        if pairs < 0:
            raise PhxError(lib().phx_last_error().decode())
        return pairs

    def set_batching(self, k1_batch=32, lines_per_batch=16):
        _check(lib().phx_lw_set_batching(self._h, int(k1_batch), int(lines_per_batch)))

    def autotune(self, k1_list=None):
        """Pick k1 batch, lines per batch and panel chunk by timing; returns them and the time."""
        k1 = None if k1_list is None else np.ascontiguousarray(k1_list, dtype=np.int32)
        out = np.zeros(4)
        _check(lib().phx_lw_autotune(self._h, 0 if k1 is None else len(k1),
                                     None if k1 is None else k1.ctypes.data_as(_ip), _p(out)))
        return dict(k1_batch=int(out[0]), lines_per_batch=int(out[1]), panel_chunk=int(out[2]),
                    seconds=float(out[3]))

    def sync(self):
        _check(lib().phx_lw_sync(self._h))

    def reset(self):
        _check(lib().phx_lw_reset(self._h))

    def set_polar(self, positions=None, born=None, dielectric=None, volume=None, coarse_mesh=None):
        """Polar term of the coupling (direct transform); no arguments: off."""
        if positions is None:
            _check(lib().phx_lw_set_polar(self._h, None))
            return
        st, self._polar_keep = polar_struct(positions, born, dielectric, self.prob.recip, volume, coarse_mesh)
        _check(lib().phx_lw_set_polar(self._h, C.byref(st)))

    def set_screen(self, nsigma):
        """Energy-conservation screen (separable transform), in units of sigma; 0 = off."""
        _check(lib().phx_lw_set_screen(self._h, float(nsigma)))

    def set_line_batching(self, min_points):
        """Batch outer points by mesh line (cacheElPh streams gW once per line) when
        a line holds at least min_points of a run's points; 0 = off."""
        _check(lib().phx_lw_set_line_batching(self._h, int(min_points)))

    def set_panel_chunk(self, panels):
        """Panels per separable-transform launch (bounds its stage buffers); 0 = automatic."""
        _check(lib().phx_lw_set_panel_chunk(self._h, int(panels)))

    def set_split_last_stage(self, on=True):
        """Separable last stage as two kernels (default) or the fused one;
        on=2 runs the split stage with the first epilogue kernel, on=3 with
        the electron rotation in the x-stage kernel (g' through device memory
        instead of g3)."""
        _check(lib().phx_lw_set_split_last_stage(self._h, on if on in (2, 3) else int(bool(on))))

    def set_epilogue_overlap(self, on=True):
        """Split last stage: each chunk's epilogue kernel on a second stream,
        concurrent with the next chunk's transform stages, or in turn (the
        handle's default)."""
        _check(lib().phx_lw_set_epilogue_overlap(self._h, int(bool(on))))

    def set_fused_y_stage(self, on=True):
        """FP32 split stage: the y-stage inside the x-stage kernel (default) or its own kernel."""
        _check(lib().phx_lw_set_fused_y_stage(self._h, int(bool(on))))

    def last_stage_fused_y(self):
        """With timing on: whether every timed split launch ran the y-stage inside the x-stage kernel."""
        v = C.c_int(0)
        _check(lib().phx_lw_last_stage_fused_y(self._h, C.byref(v)))
        return bool(v.value)

    def split_times(self):
        """Split last stage with timing on: x-stage ms, epilogue ms, launches,
        summed band counts of the rows the epilogue computed."""
        o = np.zeros(4)
        _check(lib().phx_lw_last_stage_split_times(self._h, _p(o)))
        return dict(xg_ms=o[0], ep_ms=o[1], launches=int(o[2]), bands=o[3])

    def split_bytes(self):
        """Split last stage with timing on: bytes the x-stage kernel moved (slabs
        read, g3 or g' written, U2 read when it rotates), bytes of its output
        the epilogue read."""
        o = np.zeros(2)
        _check(lib().phx_lw_last_stage_split_bytes(self._h, _p(o)))
        return dict(xg_bytes=o[0], ep_bytes=o[1])

    # ---- learned q-block screen
    def screen_collect(self, k1_list=None, sample_frac=0.02, seed=1):
        """Exact contributions of a random sample of q-blocks for training; returns the sample count."""
        k1 = None if k1_list is None else np.ascontiguousarray(k1_list, dtype=np.int32)
        n = lib().phx_lw_screen_collect(self._h, 0 if k1 is None else len(k1),
                                        None if k1 is None else k1.ctypes.data_as(_ip),
                                        float(sample_frac), int(seed))
        if n < 0:
            raise PhxError(lib().phx_last_error().decode())
        return n

    def screen_train(self, epochs=200, batch=512, lr=2e-3, seed=1):
        """Train the MLP on the GPU; returns (loss before, loss after) on the standardized targets."""
        out = np.zeros(2)
        _check(lib().phx_lw_screen_train(self._h, int(epochs), int(batch), float(lr), int(seed), _p(out)))
        return float(out[0]), float(out[1])

    def set_learned_screen(self, on=True, threshold=1e-8, audit_frac=0.01, seed=7):
        _check(lib().phx_lw_set_learned_screen(self._h, int(bool(on)), float(threshold),
                                               float(audit_frac), int(seed)))

    def screen_stats(self):
        out = np.zeros(8)
        _check(lib().phx_lw_screen_stats(self._h, _p(out)))
        keys = ("blocks_in_screen", "learned_skips", "audited", "fallbacks", "audit_max_ratio",
                "audit_mean_ratio", "train_samples", "margin")
        return {k: (int(v) if i < 4 or i == 6 else float(v)) for i, (k, v) in enumerate(zip(keys, out))}

    def screen_reset_stats(self):
        _check(lib().phx_lw_screen_reset_stats(self._h))

    def screen_params(self, params=None):
        """Return the trained network's parameters, or install `params`."""
        n = lib().phx_mlp_num_params()
        if params is None:
            out = np.zeros(n, dtype=np.float32)
            _check(lib().phx_lw_screen_params(self._h, out.ctypes.data_as(C.c_void_p), 0))
            return out
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == n
        _check(lib().phx_lw_screen_params(self._h, p.ctypes.data_as(C.c_void_p), 1))

    def set_timing(self, on=True):
        _check(lib().phx_lw_set_timing(self._h, int(bool(on))))

    def stage_times(self):
        """Last separable stage: dict(ms, launches, rows, flops)."""
        out = np.zeros(4)
        _check(lib().phx_lw_stage_times(self._h, _p(out)))
        return dict(ms=out[0], launches=int(out[1]), rows=int(out[2]), flops=out[3])

    def kernel_times(self):
        """dict: cache_ms, coupling_ms, collect_ms, batches, coupling_flops."""
        out = np.zeros(5)
        _check(lib().phx_lw_kernel_times(self._h, _p(out)))
        return dict(cache_ms=out[0], coupling_ms=out[1], collect_ms=out[2],
                    batches=int(out[3]), coupling_flops=out[4])

    def stream(self):
        return lib().phx_lw_stream(self._h)

    def linewidths(self, average=True):
        nc = len(self.prob.temperatures)
        out = np.zeros(nc * self.num_states)
        _check(lib().phx_lw_get_linewidths(self._h, _p(out), int(bool(average))))
        return out.reshape(nc, self.num_states)

    def copy_gw(self):
        p = self.prob
        out = np.empty(2 * p.nRe * p.nRp * p.nm * p.nw * p.nw)
        _check(lib().phx_lw_copy_gw(self._h, _p(out)))
        return out


def linewidths_form(prob, form=CHOSEN, gW=None, k1_batch=32, elph_file=None):
    """ElectronLinewidths set up in one of the named FORMS."""
    precision, transform, screen, line_min = FORMS[form]
    lw = ElectronLinewidths(prob, gW=gW, precision=precision, k1_batch=k1_batch, transform=transform,
                            elph_file=elph_file)
    if screen:
        lw.set_screen(screen)
    if line_min:
        lw.set_line_batching(line_min)
    return lw


def linewidths_auto(prob, gW=None, line_batching=4):
    """The reduced-precision builder for any system: the FP32 separable
    transform with the 6-sigma screen where its kernels take the shape (at
    most 32 modes, 8 windowed bands per point, nm nw <= 192), else the FP32
    direct transform; mesh-line cacheElPh for lines of >= line_batching
    points."""
    lw = None
    if prob.nm * prob.nw <= 192:
        try:
            lw = ElectronLinewidths(prob, gW=gW, precision="fp32", transform="separable")
            lw.set_screen(6.0)
        except PhxError:
            lw = None
    if lw is None:
        lw = ElectronLinewidths(prob, gW=gW, precision="fp32", transform="direct")
    if line_batching:
        lw.set_line_batching(line_batching)
    return lw


def synth_gw_host(prob) -> np.ndarray:
    """Materialize the synthetic gW on the host with the library's host twin."""
    out = np.empty(2 * prob.nRe * prob.nRp * prob.nm * prob.nw * prob.nw)
    _check(lib().phx_synth_gw_host(_p(out), prob.nw, prob.nm, prob.nRp, prob.nRe,
                                   _p(_f64(prob.gw_decay)), int(prob.gw_seed)))
    return out


# ------------------------------------------------------------------ phonon linewidths
class PhononElectronLinewidths:
    """Phonon linewidths from el-ph scattering on the GPU
    (PhElScatteringMatrix::builder, phel_scattering.cpp:43-363).

    prob: a problem object (phoebe_amd.synth.Problem or
    harmonic.LinewidthProblem) for the electrons (k mesh, window, T, mu,
    smearing) and the phonon model; coupling: an InteractionElPhWan;
    q_mesh: the phonon mesh.
    """

    def __init__(self, prob, coupling, q_mesh):
        self.prob = prob
        self.coupling = coupling          # kept alive with the handle
        self.q_mesh = tuple(int(x) for x in q_mesh)
        self._s, self._keep = make_problem_struct(prob)
        self._h = C.c_void_p()
        qm = np.array(self.q_mesh, dtype=np.int32)
        _check(lib().phx_phel_setup(C.byref(self._h), C.byref(self._s), coupling._h,
                                    qm.ctypes.data_as(_ip)))
        self.num_active = lib().phx_phel_num_active(self._h)
        self.nq = int(np.prod(self.q_mesh))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().phx_phel_destroy(self._h)
            self._h = C.c_void_p()

    def run(self, k1_list=None):
        if k1_list is None:
            pairs = lib().phx_phel_run(self._h, -1, None)
        else:
            k = np.ascontiguousarray(k1_list, dtype=np.int32)
            pairs = lib().phx_phel_run(self._h, len(k), k.ctypes.data_as(_ip))
        if pairs < 0:
            raise PhxError(lib().phx_last_error().decode())
        return pairs

    def linewidths(self, average=True):
        nc = len(self.prob.temperatures)
        out = np.zeros(nc * self.nq * self.prob.nm)
        _check(lib().phx_phel_get_linewidths(self._h, _p(out), int(bool(average))))
        return out.reshape(nc, self.nq, self.prob.nm)

    def phonons(self):
        out = np.zeros(self.nq * self.prob.nm)
        _check(lib().phx_phel_get_phonons(self._h, _p(out)))
        return out.reshape(self.nq, self.prob.nm)

    def reset(self):
        _check(lib().phx_phel_reset(self._h))

    def set_timing(self, on=True):
        _check(lib().phx_phel_set_timing(self._h, int(bool(on))))

    def kernel_times(self):
        """Totals since timing was switched on: cacheElPh / coupling /
        symmetrize + rates ms, launch counts, coupling transform FLOPs."""
        o = np.zeros(6)
        _check(lib().phx_phel_kernel_times(self._h, _p(o)))
        return dict(cache_ms=o[0], coupling_ms=o[1], reduce_ms=o[2], coupling_launches=int(o[3]),
                    cache_launches=int(o[4]), coupling_flops=o[5])
