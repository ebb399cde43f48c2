"""Small problems for the phonon-linewidth (ph-el) tests."""
import dataclasses

import numpy as np


@dataclasses.dataclass
class ToyProblem:
    """One s band on a simple cubic lattice (cosine dispersion), one atom with
    three flat modes (on-site force constants only), on-site coupling only:
    |g(k1, k1 + q, nu)|^2 = |gW_nu|^2 / M exactly, so the ph-el linewidth is a
    sum that numpy evaluates from the band formula alone.  The chemical
    potential sits off the band centre: there the cosine band's particle-hole
    symmetry would make the absorption and emission sums equal."""

    a: float
    e0: float
    t: float
    springs: np.ndarray   # 3 on-site force constants (Ry / Bohr^2)
    mass: float
    gw: np.ndarray        # 3 complex on-site couplings (Ry / Bohr)
    mesh: tuple
    temperatures: np.ndarray
    chem_pots: np.ndarray
    sigma: float
    win_min: float = -1e3
    win_max: float = 1e3
    smearing: int = 0
    name: str = "toy"

    def __post_init__(self):
        a = self.a
        self.A = a * np.eye(3)
        self.recip = 2 * np.pi / a * np.eye(3)
        self.nw = 1
        self.Rh = np.array([[0, 0, 0], [a, 0, 0], [-a, 0, 0], [0, a, 0], [0, -a, 0], [0, 0, a], [0, 0, -a]],
                           dtype=float)
        self.degh = np.ones(7)
        self.h0R = np.array([self.e0] + [self.t] * 6, dtype=complex).reshape(7, 1, 1)
        self.nm = 3
        self.Rf = np.zeros((1, 3))
        self.wf = np.ones(1)
        self.mat2R = np.diag(self.springs).reshape(1, 3, 3).astype(float)
        self.masses = np.full(3, self.mass)
        self.Re = np.zeros((1, 3))
        self.degE = np.ones(1)
        self.Rp = np.zeros((1, 3))
        self.degP = np.ones(1)
        self.long_range = None
        self.gw_decay = None
        self.gw_seed = 0

    @property
    def nRe(self):
        return 1

    @property
    def nRp(self):
        return 1

    def gw_host(self):
        return np.asarray(self.gw, dtype=complex).reshape(1, 1, 3, 1, 1)

    def band(self, k):
        a = self.a
        return self.e0 + 2 * self.t * (np.cos(k[..., 0] * a) + np.cos(k[..., 1] * a) + np.cos(k[..., 2] * a))

    def expected_phel(self, q_mesh, fold=False, sign=1.0, with_T=True):
        """Gamma(c, q, nu) = sum_k |gW_nu|^2 / M f(1-f)(e_k) delta(e_k - e_{k+q} + w_nu) pi / (T N_k)
        with GaussianDeltaFunction's one-sided cut, then the degenerate-mode
        average of row 0."""
        m = np.array(self.mesh)
        ks = np.array([[i, j, k] for k in range(m[2]) for j in range(m[1]) for i in range(m[0])]) / m
        kc = ks @ self.recip.T
        qm = np.array(q_mesh)
        qs = np.array([[i, j, k] for k in range(qm[2]) for j in range(qm[1]) for i in range(qm[0])]) / qm
        qc = qs @ self.recip.T
        w = np.sqrt(self.springs / self.mass)
        g2 = np.abs(self.gw) ** 2 / self.mass
        # degenerate modes share the averaged coupling (symmetrizeCoupling)
        g2s = g2.copy()
        for i in range(3):
            same = np.abs(w - w[i]) < 1e-6
            g2s[i] = g2[same].mean()
        e1 = self.band(kc)
        nk = len(kc)
        kmesh_pts = ks
        out = np.zeros((len(self.temperatures), len(qc), 3))
        for c, (T, mu) in enumerate(zip(self.temperatures, self.chem_pots)):
            F = 0.25 / np.cosh((e1 - mu) / (2 * T)) ** 2
            for iq, q in enumerate(qc):
                if fold:   # a planted mistake: k + q rounded onto the k mesh
                    k2 = np.round((kmesh_pts + qs[iq]) * m) / m
                    e2 = self.band(k2 @ self.recip.T)
                else:
                    e2 = self.band(kc + q)
                for nu in range(3):
                    x = (e1 - e2 + sign * w[nu]) / self.sigma
                    d = np.where(x > 6, 0.0, np.exp(-x * x) / self.sigma / np.sqrt(np.pi))
                    out[c, iq, nu] = np.sum(g2s[nu] * F * d) * np.pi / ((T if with_T else 1.0) * nk)
        for iq in range(len(qc)):
            for nu in range(3):
                same = np.abs(w - w[nu]) < 1e-6
                if same.sum() > 1:
                    out[0, iq, nu] = out[0, iq, same].sum() / same.sum()
        return out


def toy(springs=(8.0, 20.0, 45.0), q_mesh=(3, 3, 3), nt=2):
    p = ToyProblem(a=6.0, e0=0.1, t=-0.05, springs=np.array(springs, dtype=float), mass=20000.0,
                   gw=np.array([0.01 + 0.004j, -0.02j, 0.015], dtype=complex), mesh=(4, 4, 4),
                   temperatures=np.array([0.002, 0.004][:nt]), chem_pots=np.array([0.14, 0.05][:nt]),
                   sigma=0.01)
    return p, q_mesh
