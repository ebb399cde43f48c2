"""Seeded synthetic inputs for the electron-phonon linewidth path.

Builds Si-like model systems with the shapes of Phoebe's el-ph inputs:
a Wannier Hamiltonian H(R) (the `_tb.dat` content, qe_input_parser.cpp:883-1014),
reordered force constants on Wigner-Seitz vectors (the PhononH0 device arrays,
phonon_h0.cpp:672-771) and the Wannier coupling g(R_e, R_ph) (the elph file
content, interaction_elph_parsing.cpp:161-526).  Units are Rydberg atomic
units throughout (energies Ry, lengths Bohr), as in Phoebe.

The coupling tensor is large (tens of GB at Nw=16), so it is never drawn with
numpy's generator: every element comes from a counter-based hash of its flat
index, which the HIP library evaluates on the device and this module (or the
library's host twin) evaluates on the CPU with bit-identical results.  The
normal deviates are Irwin-Hall sums of four 53-bit uniforms scaled by sqrt(3)
(mean 0, variance 1), which use only exactly rounded operations.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Physical constants and conversions, restated from constants/constants.h:9-57.
PI = 3.14159265358979323846
ELECTRON_VOLT_SI = 1.6021766208e-19
HARTREE_SI = 4.359744650e-18
RYDBERG_SI = HARTREE_SI / 2.0
ENERGY_RY_TO_EV = RYDBERG_SI / ELECTRON_VOLT_SI
K_BOLTZMANN_SI = 1.38064852e-23
K_BOLTZMANN_RY = K_BOLTZMANN_SI / RYDBERG_SI
TEMPERATURE_AU_TO_SI = 1.0 / K_BOLTZMANN_RY
BOHR_RADIUS_SI = 0.52917721067e-10
DISTANCE_BOHR_TO_ANG = BOHR_RADIUS_SI / 1.0e-10
AMU_SI = 1.660538782e-27
ELECTRON_MASS_SI = 9.10938215e-31
MASS_AMU_TO_RY = AMU_SI / ELECTRON_MASS_SI / 2.0

GW_HASH_MUL = 0x9E3779B97F4A7C15
SQRT3 = 1.7320508075688772


def reciprocal_cell(A: np.ndarray) -> np.ndarray:
    """2*pi*A^-T (crystal.cpp:291); columns are the reciprocal vectors."""
    return 2.0 * PI * np.linalg.inv(A).T


def ws_vectors(grid, A: np.ndarray, super_cell_factor: int = 2):
    """Wigner-Seitz lattice vectors and degeneracies of a `grid` supercell.

    Follows Crystal::buildWignerSeitzVectors (crystal.cpp:354-459): vectors
    of a (2*nx*grid+1)^3 box are kept when they are (one of) the shortest
    among their supercell images, the degeneracy counts the equally short
    images, and the origin is swapped to index 0.  Returns (R[n,3] Cartesian,
    deg[n]).
    """
    nx = super_cell_factor
    g = np.asarray(grid, dtype=np.int64)
    i0 = np.arange(-nx * g[0], nx * g[0] + 1)
    i1 = np.arange(-nx * g[1], nx * g[1] + 1)
    i2 = np.arange(-nx * g[2], nx * g[2] + 1)
    I = np.stack(np.meshgrid(i0, i1, i2, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    j = np.arange(-nx - 1, nx + 2)
    J = np.stack(np.meshgrid(j, j, j, indexing="ij"), -1).reshape(-1, 3)
    mid = len(J) // 2
    shifts = (J * g).astype(np.float64)
    keep_vec, keep_deg = [], []
    chunk = 8192
    for s in range(0, len(I), chunk):
        d = I[s:s + chunk, None, :] - shifts[None, :, :]
        c0 = A[0, 0] * d[..., 0] + A[0, 1] * d[..., 1] + A[0, 2] * d[..., 2]
        c1 = A[1, 0] * d[..., 0] + A[1, 1] * d[..., 1] + A[1, 2] * d[..., 2]
        c2 = A[2, 0] * d[..., 0] + A[2, 1] * d[..., 1] + A[2, 2] * d[..., 2]
        dist = np.sqrt(c0 * c0 + c1 * c1 + c2 * c2)
        dmin = dist.min(axis=1)
        sel = np.abs(dist[:, mid] - dmin) < 1.0e-6
        deg = (np.abs(dist - dmin[:, None]) < 1.0e-6).sum(axis=1)
        keep_vec.append(I[s:s + chunk][sel])
        keep_deg.append(deg[sel])
    V = np.concatenate(keep_vec)
    deg = np.concatenate(keep_deg).astype(np.float64)
    tot = (1.0 / deg).sum()
    if abs(tot - float(np.prod(g))) > 1.0e-6:
        raise RuntimeError("Wigner-Seitz completeness check failed")
    R = np.stack([A[i, 0] * V[:, 0] + A[i, 1] * V[:, 1] + A[i, 2] * V[:, 2] for i in range(3)], -1)
    origin = np.nonzero(np.sqrt((V * V).sum(1)) < 1.0e-6)[0]
    o = int(origin[-1])
    R[[0, o]] = R[[o, 0]]
    deg[[0, o]] = deg[[o, 0]]
    return np.ascontiguousarray(R), deg


def minus_index(R: np.ndarray) -> np.ndarray:
    """Index of -R for every R in a Wigner-Seitz set (inversion-symmetric)."""
    key = {tuple(np.round(r, 6)): i for i, r in enumerate(R)}
    out = np.empty(len(R), dtype=np.int64)
    for i, r in enumerate(R):
        out[i] = key[tuple(np.round(-r, 6) + 0.0)]
    return out


# ----------------------------------------------------------------- hashing
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(x: np.ndarray) -> np.ndarray:
    x = x ^ (x >> np.uint64(30))
    x = x * _M1
    x = x ^ (x >> np.uint64(27))
    x = x * _M2
    x = x ^ (x >> np.uint64(31))
    return x


def hash_uniform(seed: int, ctr: np.ndarray) -> np.ndarray:
    """53-bit uniform in [0,1) from a counter: mix64(seed + ctr*golden)."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + ctr.astype(np.uint64) * np.uint64(GW_HASH_MUL)
        h = _mix64(x)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def hash_normal(seed: int, idx: np.ndarray) -> np.ndarray:
    """Irwin-Hall(4) * sqrt(3) normal deviate of counter idx (exact ops only)."""
    c = idx.astype(np.uint64) * np.uint64(4)
    s = hash_uniform(seed, c) + hash_uniform(seed, c + np.uint64(1))
    s = s + hash_uniform(seed, c + np.uint64(2))
    s = s + hash_uniform(seed, c + np.uint64(3))
    return (s - 2.0) * SQRT3


@dataclasses.dataclass
class Problem:
    """One synthetic el-ph linewidth problem (arrays in the C-ABI layouts)."""

    name: str
    seed: int
    A: np.ndarray                # direct cell, columns = lattice vectors (Bohr)
    mesh: tuple                  # Gamma-centred k mesh = q mesh
    nw: int
    Rh: np.ndarray               # (nRh,3) Cartesian
    degh: np.ndarray             # (nRh,)
    h0R: np.ndarray              # (nRh,nw,nw) complex128, Ry
    nm: int
    Rf: np.ndarray               # (nRf,3)
    wf: np.ndarray               # (nRf,)
    mat2R: np.ndarray            # (nRf,nm,nm) real, Ry/Bohr^2
    masses: np.ndarray           # (nm,) Ry mass units
    Re: np.ndarray               # (nRe,3)
    degE: np.ndarray
    Rp: np.ndarray               # (nRp,3)
    degP: np.ndarray
    gw_decay: np.ndarray         # (nRe,nRp): scale * exp(-(|Re|+|Rp|)/l), Ry
    gw_seed: int
    temperatures: np.ndarray     # Ry
    chem_pots: np.ndarray        # Ry
    sigma: float                 # Ry
    win_min: float               # Ry
    win_max: float               # Ry

    @property
    def recip(self) -> np.ndarray:
        return reciprocal_cell(self.A)

    @property
    def nRe(self) -> int:
        return len(self.Re)

    @property
    def nRp(self) -> int:
        return len(self.Rp)

    def gw_bytes(self) -> int:
        return 16 * self.nRe * self.nRp * self.nm * self.nw * self.nw

    def gw_host(self) -> np.ndarray:
        """Materialize gW as complex128 (nRe,nRp,nm,nw,nw) (small configs)."""
        per = self.nRp * self.nm * self.nw * self.nw
        out = np.empty(self.nRe * per * 2, dtype=np.float64)
        blk = self.nm * self.nw * self.nw
        for iRe in range(self.nRe):
            f = np.arange(iRe * per, (iRe + 1) * per, dtype=np.uint64)
            re = hash_normal(self.gw_seed, f * np.uint64(2))
            im = hash_normal(self.gw_seed, f * np.uint64(2) + np.uint64(1))
            d = np.repeat(self.gw_decay[iRe], blk)
            out[2 * iRe * per:2 * (iRe + 1) * per:2] = d * re
            out[2 * iRe * per + 1:2 * (iRe + 1) * per:2] = d * im
        return out.view(np.complex128).reshape(self.nRe, self.nRp, self.nm, self.nw, self.nw)


def _fcc(alat_bohr: float) -> np.ndarray:
    return 0.5 * alat_bohr * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]).T


def _hermitian_h0r(rng, R, deg, nw, length_bohr):
    nR = len(R)
    norms = np.linalg.norm(R, axis=1)
    decay = np.exp(-norms / length_bohr)
    H = (rng.standard_normal((nR, nw, nw)) + 1j * rng.standard_normal((nR, nw, nw))) / math.sqrt(2.0)
    H *= decay[:, None, None] / ENERGY_RY_TO_EV
    mi = minus_index(R)
    Hs = 0.5 * (H + np.conj(np.transpose(H[mi], (0, 2, 1))))
    return np.ascontiguousarray(Hs)


def _force_constants(rng, R, natoms, tau, length_bohr):
    """Random PD on-site block plus decaying inter-cell springs, Phi(-R)=Phi(R)^T."""
    nm = 3 * natoms
    nR = len(R)
    mat = np.zeros((nR, nm, nm))
    for iR in range(1, nR):
        for a in range(natoms):
            for b in range(natoms):
                dist = np.linalg.norm(R[iR] + tau[b] - tau[a])
                blk = rng.standard_normal((3, 3)) * 0.02 * math.exp(-dist / length_bohr)
                mat[iR, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
    mi = minus_index(R)
    mat = 0.5 * (mat + np.transpose(mat[mi], (0, 2, 1)))
    Q, _ = np.linalg.qr(rng.standard_normal((nm, nm)))
    lam = rng.uniform(0.3, 0.6, nm)
    onsite = (Q * lam) @ Q.T
    mat[0] = 0.5 * (onsite + onsite.T)
    return mat


def _midband_mu(h0R, Rh, degh, recip, nw):
    """Chemical potential at mid-band: median eigenvalue over a 4^3 sampling."""
    ev = []
    for i in range(4):
        for j in range(4):
            for k in range(4):
                kc = recip @ np.array([i / 4, j / 4, k / 4])
                ph = np.exp(1j * (Rh @ kc)) / degh
                H = np.tensordot(ph, h0R, axes=(0, 0))
                ev.append(np.linalg.eigvalsh(H))
    return round(float(np.median(np.concatenate(ev))), 6)


CONFIGS = {
    # name: (natoms, nw, el R grid, ph R grid, mesh, seed)
    "cfg0": dict(natoms=2, nw=8, rgrid_e=7, rgrid_p=6, mesh=12, seed=1000),
    "cfg2": dict(natoms=4, nw=16, rgrid_e=10, rgrid_p=9, mesh=40, seed=1002),
    "cfg4": dict(natoms=6, nw=16, rgrid_e=10, rgrid_p=9, mesh=60, seed=1004),
}


def make_problem(natoms: int, nw: int, rgrid_e: int, rgrid_p: int, mesh, seed: int,
                 name: str = "custom", temperature_k: float = 300.0,
                 sigma_ev: float = 0.01, window_ev: float = 0.5) -> Problem:
    """Si-like synthetic system with the distributions of the el-ph configs.

    * fcc cell a = 10.26 Bohr, `natoms` atoms of mass 28.0855 amu;
    * H(R) on WS(rgrid_e^3): complex N(0,1) eV * exp(-|R|/2 A), Hermitian-symmetrised;
    * force constants on WS(rgrid_p^3) with weights 1/d: random PD on-site block
      (eigenvalues U(0.3,0.6) Ry/Bohr^2) plus springs 0.02 N(0,1) exp(-|R+tb-ta|/1.5 A);
    * g(Re,Rp) on WS(rgrid_e^3) x WS(rgrid_p^3): complex N(0,1) * 0.05 eV *
      exp(-(|Re|+|Rp|)/3 A);
    * T = 300 K, mu at mid-band, Gaussian sigma 0.01 eV, window mu +- 0.5 eV.
    """
    mesh = tuple(int(m) for m in (mesh if np.iterable(mesh) else (mesh,) * 3))
    rng = np.random.default_rng(seed)
    A = _fcc(10.26)
    recip = reciprocal_cell(A)
    Re, degE = ws_vectors((rgrid_e,) * 3, A)
    Rp, degP = ws_vectors((rgrid_p,) * 3, A)
    ang = 1.0 / DISTANCE_BOHR_TO_ANG
    h0R = _hermitian_h0r(rng, Re, degE, nw, 2.0 * ang)
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25], [0.5, 0.0, 0.0], [0.75, 0.25, 0.25],
                     [0.0, 0.5, 0.0], [0.25, 0.75, 0.25], [0.0, 0.0, 0.5], [0.25, 0.25, 0.75],
                     [0.5, 0.5, 0.0], [0.75, 0.75, 0.25], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    tau = (A @ frac[:natoms].T).T
    mat2R = _force_constants(rng, Rp, natoms, tau, 1.5 * ang)
    masses = np.full(3 * natoms, 28.0855 * MASS_AMU_TO_RY)
    nRe, nRp = len(Re), len(Rp)
    decay = (0.05 / ENERGY_RY_TO_EV) * np.exp(
        -(np.linalg.norm(Re, axis=1)[:, None] + np.linalg.norm(Rp, axis=1)[None, :]) / (3.0 * ang))
    mu = _midband_mu(h0R, Re, degE, recip, nw)
    half = window_ev / ENERGY_RY_TO_EV
    return Problem(
        name=name, seed=seed, A=A, mesh=mesh, nw=nw, Rh=Re, degh=degE, h0R=h0R,
        nm=3 * natoms, Rf=Rp, wf=1.0 / degP, mat2R=np.ascontiguousarray(mat2R), masses=masses,
        Re=Re, degE=degE, Rp=Rp, degP=degP, gw_decay=np.ascontiguousarray(decay),
        gw_seed=int(rng.integers(1, 2**62)),
        temperatures=np.array([temperature_k / TEMPERATURE_AU_TO_SI]),
        chem_pots=np.array([mu]), sigma=sigma_ev / ENERGY_RY_TO_EV,
        win_min=mu - half, win_max=mu + half)


def config(name: str, **overrides) -> Problem:
    kw = dict(CONFIGS[name])
    kw.update(overrides)
    return make_problem(name=name, **kw)


def eig_family(kind: str, n: int, seed: int = 1001, rgrid: int = 8):
    """Inputs of the batched eigensolve microbenchmark: an fcc cell with R on
    WS(rgrid^3), and either H(R) for n Wannier functions ("el", same
    distribution as the el-ph configs) or force constants for n = 3 * natoms
    modes ("ph", natoms <= 12).  Returns dict(recip, R, weight, coef, masses)."""
    rng = np.random.default_rng(seed + 7919 * n + (0 if kind == "el" else 104729))
    A = _fcc(10.26)
    R, deg = ws_vectors((rgrid,) * 3, A)
    ang = 1.0 / DISTANCE_BOHR_TO_ANG
    if kind == "el":
        return dict(recip=reciprocal_cell(A), R=R, weight=deg,
                    coef=_hermitian_h0r(rng, R, deg, n, 2.0 * ang), masses=None)
    natoms = n // 3
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25], [0.5, 0.0, 0.0], [0.75, 0.25, 0.25],
                     [0.0, 0.5, 0.0], [0.25, 0.75, 0.25], [0.0, 0.0, 0.5], [0.25, 0.25, 0.75],
                     [0.5, 0.5, 0.0], [0.75, 0.75, 0.25], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    tau = (A @ frac[:natoms].T).T
    mat = _force_constants(rng, R, natoms, tau, 1.5 * ang)
    return dict(recip=reciprocal_cell(A), R=R, weight=1.0 / deg, coef=np.ascontiguousarray(mat),
                masses=np.full(n, 28.0855 * MASS_AMU_TO_RY))


_WS_COUNT = {}


def _ws_count(grid: int) -> int:
    if grid not in _WS_COUNT:
        _WS_COUNT[grid] = len(ws_vectors((grid,) * 3, _fcc(10.26))[0])
    return _WS_COUNT[grid]


def uneven_materials(count: int, seed: int = 1003, max_gw_bytes: float = 1.0e9,
                     max_mesh: int = 24):
    """Materials of the uneven-batching workload: Nw uniform in 4..32, atoms in
    1..12 (3..36 modes), R shells uniform in 5^3..9^3 (electron and phonon
    drawn separately), cubic k/q meshes 8^3..max_mesh^3, config-0 distributions
    (make_problem).  Draws whose FP32 coupling tensor exceeds max_gw_bytes are
    redrawn (a per-material device-memory budget), so the sample is the
    distribution conditioned on that size."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        nw = int(rng.integers(4, 33))
        natoms = int(rng.integers(1, 13))
        re, rp = int(rng.integers(5, 10)), int(rng.integers(5, 10))
        mesh = int(rng.integers(8, max_mesh + 1))
        mseed = int(rng.integers(1, 2**31))
        if _ws_count(re) * _ws_count(rp) * 3 * natoms * nw * nw * 8 > max_gw_bytes:
            continue
        out.append(make_problem(natoms, nw, re, rp, mesh, seed=mseed, name=f"uneven{len(out)}"))
    return out
