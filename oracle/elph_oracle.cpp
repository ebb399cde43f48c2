// Oracle: CPU FP64 restatement of Phoebe's electron-phonon linewidth path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (phoebe_amd/, include/)
// links, loads or calls this file.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, as the checker / CPU baseline.
//
// Every function restates the algorithm of the reference file:line it cites,
// keeping the reference's loop structure (GEMV over R_e per k1, separate
// g3 / g4 / gFinal loops, per-matrix eigensolves, serial linewidth loop) so
// that it can double as the "restatement" CPU baseline.
//
// Conventions (Rydberg atomic units; complex numbers interleaved re,im):
//   * energies in Ry, lengths in Bohr, wavevectors Cartesian incl. 2*pi
//     (reciprocal cell 2*pi*A^-T, crystal.cpp:291);
//   * H(k)   = sum_R e^{+ik.R}/d(R) h0R(R,m,n)      (electron_h0_wannier.cpp:803-821)
//   * D(q)   = sum_R e^{-iq.R} w(R) mat2R(R,m,n)     (phonon_h0_kokkos.cpp:23-46)
//   * g      : R_e phase e^{+ik1.Re}/de, R_ph phase e^{+iq.Rp}/dp
//              (interaction_elph.cpp:566-573, 357-366)
//   * U(k) plain, U(k+q) conjugated, e(q) plain (interaction_elph.cpp:330,387,399,642)
//   * matrices of eigenvectors are column-major (column = band / mode)
//
// Pinning (tests/test_oracle_*.py): zheev against LAPACK (scipy) and closed
// forms; Wigner-Seitz completeness sum(1/d) = prod(grid); coupling against a
// brute-force 5-fold sum and against textbook reductions (single R,
// gauge invariance, k -> k+G invariance); Gaussian/Fermi/Bose closed forms;
// linewidth reduction against hand-evaluable single-band/single-mode cases.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

using cd = std::complex<double>;

namespace {

// constants.h:9-57 (restated values)
const double kPi = 3.14159265358979323846;
const double kSqrtPi = std::sqrt(3.14159265358979323846);
const double kTwoPi = 2. * kPi;
const double kSpeedLightSi = 299792458;
const double kHPlanckSi = 6.62607015e-34;
const double kHartreeSi = 4.359744650e-18;
const double kTeraHertzAu = kHPlanckSi / kTwoPi / kHartreeSi * 1.0e+12;
const double kRyToTHz = 1. / kTeraHertzAu / 2. / kTwoPi;
const double kRyToCmm1 = 1.0e10 * kRyToTHz / kSpeedLightSi;

inline int imod(int a, int b) { return (a % b + b) % b; }  // utilities.cpp:9

inline cd ld(const double* p, size_t i) { return cd(p[2 * i], p[2 * i + 1]); }
inline void st(double* p, size_t i, cd v) {
  p[2 * i] = v.real();
  p[2 * i + 1] = v.imag();
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Wigner-Seitz lattice vectors and degeneracies: crystal.cpp:354-459.
// A is the direct cell, column-major (A[i + 3*j] = component i of vector j).
// Returns the number of vectors; when R/deg are non-null (capacity maxOut)
// they are filled with Cartesian vectors (row-major n x 3) and degeneracies,
// with the origin swapped to index 0 (crystal.cpp:447-456).
// Returns -1 if the completeness check fails, -2 if maxOut is too small.
// ---------------------------------------------------------------------------
int orc_ws_vectors(const int* grid, int superCellFactor, const double* A,
                   int maxOut, double* R, double* deg) {
  const int nx = superCellFactor;
  int midIndex = 1;
  for (int d = 0; d < 3; ++d) midIndex *= 2 * (nx + 1) + 1;
  midIndex /= 2;
  std::vector<double> vecs;
  std::vector<int> degs;
  std::vector<double> distances;
  for (int i0 = -nx * grid[0]; i0 <= nx * grid[0]; i0++) {
    for (int i1 = -nx * grid[1]; i1 <= nx * grid[1]; i1++) {
      for (int i2 = -nx * grid[2]; i2 <= nx * grid[2]; i2++) {
        distances.clear();
        for (int j0 = -nx - 1; j0 <= nx + 1; j0++) {
          for (int j1 = -nx - 1; j1 <= nx + 1; j1++) {
            for (int j2 = -nx - 1; j2 <= nx + 1; j2++) {
              double d0 = i0 - j0 * grid[0];
              double d1 = i1 - j1 * grid[1];
              double d2 = i2 - j2 * grid[2];
              double c[3];
              for (int i = 0; i < 3; ++i)
                c[i] = A[i + 0] * d0 + A[i + 3] * d1 + A[i + 6] * d2;
              distances.push_back(std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
            }
          }
        }
        double distMin = *std::min_element(distances.begin(), distances.end());
        if (std::abs(distances[midIndex] - distMin) < 1.0e-6) {
          int degeneracy = 0;
          for (double dd : distances)
            if (std::abs(dd - distMin) < 1.0e-6) ++degeneracy;
          degs.push_back(degeneracy);
          vecs.push_back(i0);
          vecs.push_back(i1);
          vecs.push_back(i2);
        }
      }
    }
  }
  double tot = 0.;
  for (int x : degs) tot += 1. / double(x);
  if (std::abs(tot - double(grid[0]) * grid[1] * grid[2]) > 1.0e-6) return -1;
  const int n = int(degs.size());
  if (R == nullptr || deg == nullptr) return n;
  if (n > maxOut) return -2;
  int originIndex = -1;
  for (int iR = 0; iR < n; ++iR) {
    const double* v = &vecs[3 * iR];
    for (int i = 0; i < 3; ++i)
      R[3 * iR + i] = A[i] * v[0] + A[i + 3] * v[1] + A[i + 6] * v[2];
    deg[iR] = degs[iR];
    if (std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < 1.0e-6) originIndex = iR;
  }
  if (originIndex < 0) return -1;
  std::swap(deg[0], deg[originIndex]);
  for (int i = 0; i < 3; ++i) std::swap(R[i], R[3 * originIndex + i]);
  return n;
}

// ---------------------------------------------------------------------------
// Mesh folding: Points::isPointStored for a Gamma-centred mesh
// (points.cpp:404-423). Note the signed sum of (round(x) - x).
// ---------------------------------------------------------------------------
int orc_mesh_index(const double* xcrys, const int* mesh) {
  int p[3];
  double diff = 0.;
  for (int i = 0; i < 3; ++i) {
    double x = xcrys[i] * mesh[i];
    diff += std::round(x) - x;
    p[i] = imod(int(std::round(x)), mesh[i]);
  }
  if (diff >= 1.0e-6) return -1;
  return p[2] * mesh[0] * mesh[1] + p[1] * mesh[0] + p[0];
}

// Mesh point -> crystal coordinates (points.cpp:233-240, offset 0).
void orc_mesh_point(int index, const int* mesh, double* xcrys) {
  int ikz = index / (mesh[0] * mesh[1]);
  int idx = index - ikz * mesh[0] * mesh[1];
  int iky = idx / mesh[0];
  int ikx = imod(idx, mesh[0]);
  xcrys[0] = double(ikx) / double(mesh[0]);
  xcrys[1] = double(iky) / double(mesh[1]);
  xcrys[2] = double(ikz) / double(mesh[2]);
}

// ---------------------------------------------------------------------------
// Hermitian eigensolver (the oracle's own, cyclic complex Jacobi).
// Replaces Eigen::SelfAdjointEigenSolver as used by kokkosZHEEV
// (common_kokkos.cpp:26-53): eigenvalues ascending, eigenvectors as columns
// of a column-major matrix written over A.  Pinned against LAPACK zheev.
// Returns the number of sweeps used, or -1 if not converged.
// ---------------------------------------------------------------------------
int orc_zheev(int N, double* Ainout, double* W) {
  std::vector<cd> a(size_t(N) * N), v(size_t(N) * N, cd(0.));
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) a[i + size_t(N) * j] = ld(Ainout, i + size_t(N) * j);
  // use the lower triangle as the reference (uplo = L) and hermitize exactly
  for (int j = 0; j < N; ++j) {
    a[j + size_t(N) * j] = cd(a[j + size_t(N) * j].real(), 0.);
    for (int i = j + 1; i < N; ++i) a[j + size_t(N) * i] = std::conj(a[i + size_t(N) * j]);
  }
  for (int i = 0; i < N; ++i) v[i + size_t(N) * i] = 1.;
  auto A = [&](int i, int j) -> cd& { return a[i + size_t(N) * j]; };
  auto V = [&](int i, int j) -> cd& { return v[i + size_t(N) * j]; };
  double norm2 = 0.;
  for (auto& x : a) norm2 += std::norm(x);
  int sweep = 0;
  const int maxSweeps = 60;
  for (; sweep < maxSweeps; ++sweep) {
    double off = 0.;
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i)
        if (i != j) off += std::norm(A(i, j));
    if (off <= 1e-30 * norm2 || off == 0.) break;
    for (int p = 0; p < N - 1; ++p) {
      for (int q = p + 1; q < N; ++q) {
        cd apq = A(p, q);
        double r = std::abs(apq);
        if (r == 0.) continue;
        double app = A(p, p).real(), aqq = A(q, q).real();
        if (r < 1e-300) continue;
        cd ph = apq / r;  // e^{i phi}
        double tau = (aqq - app) / (2. * r);
        double t = (tau >= 0. ? 1. : -1.) / (std::abs(tau) + std::sqrt(1. + tau * tau));
        double c = 1. / std::sqrt(1. + t * t);
        double s = t * c;
        // J = [[c, s e^{i phi}], [-s e^{-i phi}, c]] on (p,q); A <- J^H A J, V <- V J
        cd sp = s * ph, sm = s * std::conj(ph);
        for (int k = 0; k < N; ++k) {  // columns
          cd akp = A(k, p), akq = A(k, q);
          A(k, p) = akp * c - akq * sm;
          A(k, q) = akp * sp + akq * c;
        }
        for (int k = 0; k < N; ++k) {  // rows
          cd apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - sp * aqk;
          A(q, k) = sm * apk + c * aqk;
        }
        A(p, q) = 0.;
        A(q, p) = 0.;
        A(p, p) = cd(A(p, p).real(), 0.);
        A(q, q) = cd(A(q, q).real(), 0.);
        for (int k = 0; k < N; ++k) {
          cd vkp = V(k, p), vkq = V(k, q);
          V(k, p) = vkp * c - vkq * sm;
          V(k, q) = vkp * sp + vkq * c;
        }
      }
    }
  }
  std::vector<int> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return A(x, x).real() < A(y, y).real(); });
  for (int j = 0; j < N; ++j) {
    W[j] = A(order[j], order[j]).real();
    for (int i = 0; i < N; ++i) st(Ainout, i + size_t(N) * j, V(i, order[j]));
  }
  return sweep < maxSweeps ? sweep : -1;
}

// ---------------------------------------------------------------------------
// Electron Bloch Hamiltonian: electron_h0_wannier.cpp:160-216 (CPU path)
// and :803-821 (Kokkos path): H(k)_{mn} = sum_R e^{i k.R}/d(R) h0R(R,m,n).
// h0R is (nR, nw, nw) row-major interleaved; H is column-major.
// ---------------------------------------------------------------------------
void orc_el_hamiltonian(int nw, int nR, const double* R, const double* deg,
                        const double* h0R, const double* k, double* H) {
  std::vector<cd> phases(nR);
  for (int iR = 0; iR < nR; ++iR) {
    double arg = k[0] * R[3 * iR] + k[1] * R[3 * iR + 1] + k[2] * R[3 * iR + 2];
    phases[iR] = cd(std::cos(arg), std::sin(arg)) / deg[iR];
  }
  for (int n = 0; n < nw; ++n)
    for (int m = 0; m < nw; ++m) {
      cd tmp = 0.;
      for (int iR = 0; iR < nR; ++iR) tmp += phases[iR] * ld(h0R, (size_t(iR) * nw + m) * nw + n);
      st(H, m + size_t(nw) * n, tmp);
    }
}

// Energies + eigenvectors at k (electron_h0_wannier.cpp:218-235 / 855-877).
int orc_el_diag(int nw, int nR, const double* R, const double* deg, const double* h0R,
                const double* k, double* energies, double* eigvecs) {
  orc_el_hamiltonian(nw, nR, R, deg, h0R, k, eigvecs);
  return orc_zheev(nw, eigvecs, energies) < 0 ? -1 : 0;
}

// ---------------------------------------------------------------------------
// Phonon dynamical matrix: phonon_h0_kokkos.cpp:6-46 (FT with e^{-iq.R} w(R)),
// :150-170 (hermitize, divide by sqrt(M_m M_n)); mat2R is (nR, nm, nm).
// ---------------------------------------------------------------------------
// Rigid-ion long-range term (PhononH0::addLongRangeTerm, phonon_h0.cpp:537-670,
// the CPU form of QE's rgd_blk): constant part on the diagonal blocks, then
// the G-space sum with k = G + q,
//   dyn(i, j, na, nb) += normG(k) e^{i (tau_na - tau_nb).k} kZ(i, na) kZ(j, nb),
//   kZ(i, nb) = sum_l k_l born(nb, l, i),  normG = norm e^{-k.eps.k / 4} / (k.eps.k)
//   for 0 < k.eps.k < 4 gMax.
struct orc_long_range {
  int nat, nG;
  const double *tau, *born, *eps, *G, *cst;  // nat x 3, nat x 9, 9, nG x 3, nat x 9
  double norm, gMax;
};

static void orc_add_long_range(int nm, const orc_long_range& lr, const double* q, cd* dyn /* (m, n) row-major */) {
  const int nat = lr.nat;
  for (int na = 0; na < nat; ++na)
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) dyn[size_t(na * 3 + i) * nm + na * 3 + j] += lr.cst[9 * na + 3 * i + j];
  std::vector<double> kZ(size_t(3) * nat);
  std::vector<cd> ph(size_t(nat) * nat);
  for (int ig = 0; ig < lr.nG; ++ig) {
    double k[3];
    for (int i = 0; i < 3; ++i) k[i] = lr.G[3 * ig + i] + q[i];
    double geg = 0.;
    for (int i = 0; i < 3; ++i) {
      double t = 0.;
      for (int j = 0; j < 3; ++j) t += lr.eps[3 * i + j] * k[j];
      geg += k[i] * t;
    }
    if (!(geg > 0. && geg < 4. * lr.gMax)) continue;
    const double normG = lr.norm * std::exp(-geg * 0.25) / geg;
    for (int i = 0; i < 3; ++i)
      for (int nb = 0; nb < nat; ++nb)
        kZ[size_t(i) * nat + nb] = k[0] * lr.born[9 * nb + 0 + i] + k[1] * lr.born[9 * nb + 3 + i] +
                                   k[2] * lr.born[9 * nb + 6 + i];
    for (int nb = 0; nb < nat; ++nb)
      for (int na = 0; na < nat; ++na) {
        double arg = 0.;
        for (int i = 0; i < 3; ++i) arg += (lr.tau[3 * na + i] - lr.tau[3 * nb + i]) * k[i];
        ph[size_t(na) * nat + nb] = cd(std::cos(arg), std::sin(arg));
      }
    for (int nb = 0; nb < nat; ++nb)
      for (int na = 0; na < nat; ++na)
        for (int j = 0; j < 3; ++j)
          for (int i = 0; i < 3; ++i)
            dyn[size_t(na * 3 + i) * nm + nb * 3 + j] +=
                normG * ph[size_t(na) * nat + nb] * cd(kZ[size_t(i) * nat + na] * kZ[size_t(j) * nat + nb], 0.);
  }
}

static void orc_ph_dynamical_lr(int nm, int nR, const double* R, const double* w, const double* mat2R,
                                const double* masses, const double* q, const orc_long_range* lr,
                                double* D);

void orc_ph_dynamical(int nm, int nR, const double* R, const double* w, const double* mat2R,
                      const double* masses, const double* q, double* D) {
  orc_ph_dynamical_lr(nm, nR, R, w, mat2R, masses, q, nullptr, D);
}

static void orc_ph_dynamical_lr(int nm, int nR, const double* R, const double* w, const double* mat2R,
                                const double* masses, const double* q, const orc_long_range* lr,
                                double* D) {
  std::vector<cd> phases(nR);
  for (int iR = 0; iR < nR; ++iR) {
    double arg = q[0] * R[3 * iR] + q[1] * R[3 * iR + 1] + q[2] * R[3 * iR + 2];
    phases[iR] = std::exp(-cd(0., 1.) * arg) * w[iR];
  }
  std::vector<cd> dyn(size_t(nm) * nm);
  for (int m = 0; m < nm; ++m)
    for (int n = 0; n < nm; ++n) {
      cd tmp = 0.;
      for (int iR = 0; iR < nR; ++iR) tmp += phases[iR] * mat2R[(size_t(iR) * nm + m) * nm + n];
      dyn[size_t(m) * nm + n] = tmp;
    }
  if (lr && lr->nG > 0) orc_add_long_range(nm, *lr, q, dyn.data());
  for (int n = 0; n < nm; ++n)
    for (int m = 0; m < nm; ++m) {
      cd x = 0.5 * (dyn[size_t(m) * nm + n] + std::conj(dyn[size_t(n) * nm + m]));
      x /= std::sqrt(masses[m] * masses[n]);
      st(D, m + size_t(nm) * n, x);
    }
}

// Frequencies omega = sign(l) sqrt|l| and mass-scaled eigenvectors e/sqrt(M)
// (phonon_h0_kokkos.cpp:177-219).
int orc_ph_diag_lr(int nm, int nR, const double* R, const double* w, const double* mat2R,
                   const double* masses, const double* q, const orc_long_range* lr, double* omega,
                   double* eigvecs) {
  orc_ph_dynamical_lr(nm, nR, R, w, mat2R, masses, q, lr, eigvecs);
  if (orc_zheev(nm, eigvecs, omega) < 0) return -1;
  for (int m = 0; m < nm; ++m)
    omega[m] = omega[m] > 0. ? std::sqrt(omega[m]) : -std::sqrt(-omega[m]);
  for (int n = 0; n < nm; ++n)
    for (int m = 0; m < nm; ++m) {
      size_t i = m + size_t(nm) * n;
      st(eigvecs, i, ld(eigvecs, i) / std::sqrt(masses[m]));
    }
  return 0;
}

int orc_ph_diag(int nm, int nR, const double* R, const double* w, const double* mat2R,
                const double* masses, const double* q, double* omega, double* eigvecs) {
  return orc_ph_diag_lr(nm, nR, R, w, mat2R, masses, q, nullptr, omega, eigvecs);
}

// ---------------------------------------------------------------------------
// InteractionElPhWan::cacheElPh (interaction_elph.cpp:492-701), pool size 1,
// CPU (non-CUDA) branch: phases (:566-573), GEMV over R_e (:604-608), U1
// rotation without conjugation (:635-645).
// gW: (nRe, nRp, nm, nw(w1), nw(w2)) row-major interleaved (the reference's
// LayoutRight view of the Eigen tensor, :44-47).
// U1: (nw, nb1) column-major.  cached: (nRp, nm, nb1, nw(w2)) row-major.
// ---------------------------------------------------------------------------
void orc_cache_elph(int nw, int nm, int nRp, int nRe, const double* gW, const double* Re,
                    const double* degE, const double* U1, int nb1, const double* k1,
                    double* cached) {
  std::vector<cd> phases(nRe);
  for (int irE = 0; irE < nRe; ++irE) {
    double arg = k1[0] * Re[3 * irE] + k1[1] * Re[3 * irE + 1] + k1[2] * Re[3 * irE + 2];
    phases[irE] = std::exp(cd(0., 1.) * arg) / degE[irE];
  }
  const size_t ncol = size_t(nRp) * nm * nw * nw;
  std::vector<cd> g1(ncol);
  // gemv("T"): g1(col) = sum_irE gW(irE, col) * phase(irE)
#pragma omp parallel for schedule(static)
  for (long long col = 0; col < (long long)ncol; ++col) {
    cd tmp = 0.;
    for (int irE = 0; irE < nRe; ++irE) tmp += ld(gW, size_t(irE) * ncol + col) * phases[irE];
    g1[col] = tmp;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int irP = 0; irP < nRp; ++irP)
    for (int nu = 0; nu < nm; ++nu)
      for (int ib1 = 0; ib1 < nb1; ++ib1)
        for (int iw2 = 0; iw2 < nw; ++iw2) {
          cd tmp = 0.;
          for (int iw1 = 0; iw1 < nw; ++iw1)
            tmp += g1[((size_t(irP) * nm + nu) * nw + iw1) * nw + iw2] *
                   ld(U1, iw1 + size_t(nw) * ib1);
          st(cached, ((size_t(irP) * nm + nu) * nb1 + ib1) * nw + iw2, tmp);
        }
}

// ---------------------------------------------------------------------------
// InteractionElPhWan::calcCouplingSquared (interaction_elph.cpp:229-448),
// polar correction off.  phases (:357-366), g3 (:369-378), g4 (:381-390),
// gFinal (:393-402), |g|^2 with the b1/b2 flip (:416-446).
// U2s: packed per q, each (nw, nb2_q) column-major.  e3s: nq x (nm, nm)
// column-major (rows = 3*atom+pol, columns = modes).  q3: nq x 3 Cartesian.
// out: packed per q, each (nb1, nb2_q, nm) column-major.
// ---------------------------------------------------------------------------
// polarCorrectionPart1Static (interaction_elph.cpp:152-203): x(nu) summed
// over q+G, G = m1 b1 + m2 b2 + m3 b3 with |m_i| <= coarse[i], for
// 0 < qEq = (q+G).eps.(q+G) and qEq / 4 < gMax = 14, of
// i chargeSquare fourPi / V e^{-qEq/4} / qEq e^{-i (q+G).tau} ((q+G).Z) ev3(k, nu),
// chargeSquare = 2, fourPi = 2 pi (constants.h:12), k = 3 iAt + iPol
// (compress2Indices, utilities.cpp:29-33).  born: (natoms, 3, 3) Z(iAt, i, iPol);
// eps (3, 3) row-major; recip column-major; e3 (nm, nm) column-major; x nm complex.
void orc_polar_part1(int natoms, const double* pos, const double* born, const double* eps,
                     const double* recip, double volume, const int* coarse, int nm, const double* q,
                     const double* e3, double* x) {
  const double gMax = 14.;
  const double chargeSquare = 2.;
  const double fourPi = 2. * kPi;
  const cd factor = chargeSquare * fourPi / volume * cd(0., 1.);
  std::vector<cd> xs(nm, cd(0.));
  for (int m1 = -coarse[0]; m1 <= coarse[0]; m1++)
    for (int m2 = -coarse[1]; m2 <= coarse[1]; m2++)
      for (int m3 = -coarse[2]; m3 <= coarse[2]; m3++) {
        double G[3];
        for (int i = 0; i < 3; ++i) G[i] = recip[i] * m1 + recip[i + 3] * m2 + recip[i + 6] * m3 + q[i];
        double qEq = 0.;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) qEq += G[i] * eps[3 * i + j] * G[j];
        if (!(qEq > 0. && qEq / 4. < gMax)) continue;
        const cd factor2 = factor * std::exp(-qEq / 4.) / qEq;
        for (int iAt = 0; iAt < natoms; iAt++) {
          const double arg = -(G[0] * pos[3 * iAt] + G[1] * pos[3 * iAt + 1] + G[2] * pos[3 * iAt + 2]);
          const cd factor3 = factor2 * cd(std::cos(arg), std::sin(arg));
          for (int iPol = 0; iPol < 3; ++iPol) {
            const double* Z = born + 9 * iAt;
            const double gqDotZ = G[0] * Z[iPol] + G[1] * Z[3 + iPol] + G[2] * Z[6 + iPol];
            const int k = 3 * iAt + iPol;
            for (int ib3 = 0; ib3 < nm; ib3++) xs[ib3] += factor3 * gqDotZ * ld(e3, k + size_t(nm) * ib3);
          }
        }
      }
  for (int ib3 = 0; ib3 < nm; ++ib3) st(x, ib3, xs[ib3]);
}

// calcCouplingSquared with the polar term: polar (nq x nm complex, or null)
// is polarCorrectionPart1 of each q; where usePolarCorrection and |q3| > 1e-8
// (:285-312) gFinal(nu, b1, b2) += x(nu) overlap(b1, b2), overlap = (U2^H U1)^T
// (polarCorrectionPart2, :205-227), before |g|^2 (:406-413).
void orc_coupling_sq_polar(int nw, int nm, int nRp, const double* Rp, const double* degP,
                           const double* cached, int nb1, int nq, const double* U2s,
                           const int* nb2s, const double* e3s, const double* q3, const double* U1,
                           const double* polar, double* out) {
  std::vector<size_t> u2off(nq + 1, 0), outoff(nq + 1, 0);
  int nb2max = 0;
  for (int iq = 0; iq < nq; ++iq) {
    u2off[iq + 1] = u2off[iq] + size_t(nw) * nb2s[iq];
    outoff[iq + 1] = outoff[iq] + size_t(nb1) * nb2s[iq] * nm;
    nb2max = std::max(nb2max, nb2s[iq]);
  }
#pragma omp parallel for schedule(dynamic)
  for (int ik = 0; ik < nq; ++ik) {
    std::vector<cd> phases(nRp);
    for (int irP = 0; irP < nRp; ++irP) {
      double arg = q3[3 * ik] * Rp[3 * irP] + q3[3 * ik + 1] * Rp[3 * irP + 1] +
                   q3[3 * ik + 2] * Rp[3 * irP + 2];
      phases[irP] = std::exp(cd(0., 1.) * arg) / degP[irP];
    }
    std::vector<cd> g3(size_t(nm) * nb1 * nw), g4(size_t(nm) * nb1 * nw);
    for (int nu = 0; nu < nm; ++nu)
      for (int ib1 = 0; ib1 < nb1; ++ib1)
        for (int iw2 = 0; iw2 < nw; ++iw2) {
          cd tmp = 0.;
          for (int irP = 0; irP < nRp; ++irP)
            tmp += phases[irP] * ld(cached, ((size_t(irP) * nm + nu) * nb1 + ib1) * nw + iw2);
          g3[(size_t(nu) * nb1 + ib1) * nw + iw2] = tmp;
        }
    const double* e3 = e3s + 2 * size_t(ik) * nm * nm;
    for (int nu2 = 0; nu2 < nm; ++nu2)
      for (int ib1 = 0; ib1 < nb1; ++ib1)
        for (int iw2 = 0; iw2 < nw; ++iw2) {
          cd tmp = 0.;
          // eigvecs3_k(ik, nu2, nu) = eigvecs3[ik](nu, nu2)   (:336-340)
          for (int nu = 0; nu < nm; ++nu)
            tmp += g3[(size_t(nu) * nb1 + ib1) * nw + iw2] * ld(e3, nu + size_t(nm) * nu2);
          g4[(size_t(nu2) * nb1 + ib1) * nw + iw2] = tmp;
        }
    const int nb2 = nb2s[ik];
    const double* U2 = U2s + 2 * u2off[ik];
    double* o = out + outoff[ik];
    const double qn = std::sqrt(q3[3 * ik] * q3[3 * ik] + q3[3 * ik + 1] * q3[3 * ik + 1] +
                                q3[3 * ik + 2] * q3[3 * ik + 2]);
    const bool usePolar = polar != nullptr && qn > 1.0e-8;
    std::vector<cd> overlap;  // (nb1, nb2)
    if (usePolar) {
      overlap.assign(size_t(nb1) * nb2, cd(0.));
      for (int ib1 = 0; ib1 < nb1; ++ib1)
        for (int ib2 = 0; ib2 < nb2; ++ib2) {
          cd t = 0.;
          for (int w = 0; w < nw; ++w)
            t += std::conj(ld(U2, w + size_t(nw) * ib2)) * ld(U1, w + size_t(nw) * ib1);
          overlap[ib1 + size_t(nb1) * ib2] = t;
        }
    }
    for (int nu = 0; nu < nm; ++nu)
      for (int ib1 = 0; ib1 < nb1; ++ib1)
        for (int ib2 = 0; ib2 < nb2; ++ib2) {
          cd tmp = 0.;
          for (int iw2 = 0; iw2 < nw; ++iw2)
            tmp += std::conj(ld(U2, iw2 + size_t(nw) * ib2)) * g4[(size_t(nu) * nb1 + ib1) * nw + iw2];
          if (usePolar) tmp += ld(polar, size_t(ik) * nm + nu) * overlap[ib1 + size_t(nb1) * ib2];
          o[ib1 + size_t(nb1) * (ib2 + size_t(nb2) * nu)] = std::norm(tmp);
        }
  }
}

void orc_coupling_sq(int nw, int nm, int nRp, const double* Rp, const double* degP,
                     const double* cached, int nb1, int nq, const double* U2s,
                     const int* nb2s, const double* e3s, const double* q3, double* out) {
  orc_coupling_sq_polar(nw, nm, nRp, Rp, degP, cached, nb1, nq, U2s, nb2s, e3s, q3, nullptr, nullptr, out);
}

// ---------------------------------------------------------------------------
// ScatteringMatrix::symmetrizeCoupling (scattering.cpp:1209-1318): average
// |g|^2 over consecutive degenerate groups (|e - e_first| < 1e-6 Ry) along
// b1, then b2, then nu.  coupling is (nb1, nb2, nm) column-major.
// ---------------------------------------------------------------------------
void orc_symmetrize_coupling(double* coupling, int nb1, int nb2, int nm, const double* e1,
                             const double* e2, const double* e3) {
  auto C = [&](int a, int b, int c) -> double& {
    return coupling[a + size_t(nb1) * (b + size_t(nb2) * c)];
  };
  std::vector<double> tmp;
  for (int ib1 = 0; ib1 < nb1; ib1++) {
    int deg = 0;
    for (int i = ib1; i < nb1; i++) {
      if (std::abs(e1[ib1] - e1[i]) < 1.0e-6) deg++;
      else break;
    }
    tmp.assign(size_t(nb2) * nm, 0.);
    for (int i = 0; i < deg; ++i)
      for (int ib2 = 0; ib2 < nb2; ib2++)
        for (int ib3 = 0; ib3 < nm; ib3++)
          tmp[ib2 + size_t(nb2) * ib3] += C(ib1 + i, ib2, ib3) / float(deg);
    for (int i = 0; i < deg; ++i)
      for (int ib2 = 0; ib2 < nb2; ib2++)
        for (int ib3 = 0; ib3 < nm; ib3++) C(ib1 + i, ib2, ib3) = tmp[ib2 + size_t(nb2) * ib3];
    ib1 += deg - 1;
  }
  for (int ib2 = 0; ib2 < nb2; ib2++) {
    int deg = 0;
    for (int i = ib2; i < nb2; i++) {
      if (std::abs(e2[ib2] - e2[i]) < 1.0e-6) deg++;
      else break;
    }
    tmp.assign(size_t(nb1) * nm, 0.);
    for (int i = 0; i < deg; ++i)
      for (int ib1 = 0; ib1 < nb1; ib1++)
        for (int ib3 = 0; ib3 < nm; ib3++)
          tmp[ib1 + size_t(nb1) * ib3] += C(ib1, ib2 + i, ib3) / float(deg);
    for (int i = 0; i < deg; ++i)
      for (int ib1 = 0; ib1 < nb1; ib1++)
        for (int ib3 = 0; ib3 < nm; ib3++) C(ib1, ib2 + i, ib3) = tmp[ib1 + size_t(nb1) * ib3];
    ib2 += deg - 1;
  }
  for (int ib3 = 0; ib3 < nm; ib3++) {
    int deg = 0;
    for (int i = ib3; i < nm; i++) {
      if (std::abs(e3[ib3] - e3[i]) < 1.0e-6) deg++;
      else break;
    }
    tmp.assign(size_t(nb1) * nb2, 0.);
    for (int i = 0; i < deg; ++i)
      for (int ib1 = 0; ib1 < nb1; ib1++)
        for (int ib2 = 0; ib2 < nb2; ib2++)
          tmp[ib1 + size_t(nb1) * ib2] += C(ib1, ib2, ib3 + i) / float(deg);
    for (int i = 0; i < deg; ++i)
      for (int ib1 = 0; ib1 < nb1; ib1++)
        for (int ib2 = 0; ib2 < nb2; ib2++) C(ib1, ib2, ib3 + i) = tmp[ib1 + size_t(nb1) * ib2];
    ib3 += deg - 1;
  }
}

// GaussianDeltaFunction::getSmearing (delta_function.cpp:34-45): one-sided cut.
double orc_gaussian_delta(double energy, double sigma) {
  double inverseWidth = 1. / sigma;
  double prefactor = 1. / sigma / std::sqrt(kPi);
  double x = energy * inverseWidth;
  if (x > 6.) return 0.;
  return prefactor * std::exp(-x * x);
}

// AdaptiveGaussianDeltaFunction (delta_function.cpp:55-97): qTensor = the
// reciprocal cell with row i divided by mesh(i) (:58-62); width
// sqrt(sum_i (qTensor.row(i) . v)^2 / 6) with v = v1 - v2, at least 0.1 meV,
// the Gaussian cut at two widths and divided by erf(2).
double orc_adaptive_delta(double energy, const double* v, const double* recip, const int* mesh) {
  if (v[0] == 0. && v[1] == 0. && v[2] == 0. && energy == 0.) return 1.;
  double sigma = 0.;
  for (int i = 0; i < 3; ++i) {
    double d = 0.;
    for (int j = 0; j < 3; ++j) d += recip[i + 3 * j] / mesh[i] * v[j];
    sigma += d * d;
  }
  sigma = std::sqrt(sigma / 6.);
  if (sigma == 0.) return 0.;
  const double energyRyToEv = (4.359744650e-18 / 2.) / 1.6021766208e-19;
  const double cutoff = 0.0001 / energyRyToEv;
  if (sigma < cutoff) sigma = cutoff;
  if (std::abs(energy) > 2. * sigma) return 0.;
  double x = energy / sigma;
  if (x > 6.) return 0.;
  return std::exp(-x * x) / std::sqrt(kPi) / sigma / 0.9953222650189527;
}

// Group velocities (the diagonal of the velocity operator) of
// ElectronH0Wannier::batchedDiagonalizeWithVelocities (electron_h0_wannier.cpp:
// 633-760): central differences (1e-8) of H(k), hermitized, rotated to the
// eigenbasis; zero when |k| < 1e-6; a degenerate group (1 micro-eV) gets the
// eigenvalues of its block, ascending.  vg: nw x 3.
int orc_el_group_velocities_rule(int nw, int nR, const double* R, const double* deg,
                                 const double* h0R, const double* k, int gammaRule, double* vg);

int orc_el_group_velocities(int nw, int nR, const double* R, const double* deg, const double* h0R,
                            const double* k, double* vg) {
  return orc_el_group_velocities_rule(nw, nR, R, deg, h0R, k, 1, vg);
}

// gammaRule 1: zero when |k| < 1e-6 (single-point form); 0: unless |k|^2 >
// 1e-6 (the batched form, electron_h0_wannier.cpp:978-983).
int orc_el_group_velocities_rule(int nw, int nR, const double* R, const double* deg,
                                 const double* h0R, const double* k, int gammaRule, double* vg) {
  const double delta = 1.0e-8;
  const double energyRyToEv = (4.359744650e-18 / 2.) / 1.6021766208e-19;
  const double threshold = 0.000001 / energyRyToEv;
  const size_t nn = size_t(nw) * nw;
  std::vector<double> U(2 * nn), E(nw);
  orc_el_hamiltonian(nw, nR, R, deg, h0R, k, U.data());
  if (orc_zheev(nw, U.data(), E.data()) < 0) return -1;
  std::fill(vg, vg + 3 * nw, 0.);
  const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
  if (gammaRule ? std::sqrt(k2) < 1.0e-6 : !(k2 > 1.0e-6)) return 0;
  std::vector<double> Hp(2 * nn), Hm(2 * nn);
  std::vector<cd> V(nn * 3);
  for (int i = 0; i < 3; ++i) {
    double kp[3] = {k[0], k[1], k[2]}, km[3] = {k[0], k[1], k[2]};
    kp[i] += delta;
    km[i] -= delta;
    orc_el_hamiltonian(nw, nR, R, deg, h0R, kp, Hp.data());
    orc_el_hamiltonian(nw, nR, R, deg, h0R, km, Hm.data());
    std::vector<cd> der(nn);
    for (int n = 0; n < nw; ++n)
      for (int m = 0; m < nw; ++m) {
        const size_t a = m + size_t(nw) * n, b = n + size_t(nw) * m;
        cd x = (cd(Hp[2 * a], Hp[2 * a + 1]) - cd(Hm[2 * a], Hm[2 * a + 1])) / (2. * delta);
        cd y = (cd(Hp[2 * b], Hp[2 * b + 1]) - cd(Hm[2 * b], Hm[2 * b + 1])) / (2. * delta);
        der[a] = 0.5 * (x + std::conj(y));
      }
    // V_i = U^H der U
    for (int n = 0; n < nw; ++n)
      for (int m = 0; m < nw; ++m) {
        cd acc = 0.;
        for (int l = 0; l < nw; ++l)
          for (int l2 = 0; l2 < nw; ++l2)
            acc += std::conj(cd(U[2 * (l + size_t(nw) * m)], U[2 * (l + size_t(nw) * m) + 1])) *
                   der[l + size_t(nw) * l2] * cd(U[2 * (l2 + size_t(nw) * n)], U[2 * (l2 + size_t(nw) * n) + 1]);
        V[(m + size_t(nw) * n) * 3 + i] = acc;
      }
  }
  for (int ib = 0; ib < nw;) {
    int sz = 1;
    while (ib + sz < nw && std::abs(E[ib] - E[ib + sz]) <= threshold) ++sz;
    for (int i = 0; i < 3; ++i) {
      if (sz == 1) {
        vg[3 * ib + i] = V[(ib + size_t(nw) * ib) * 3 + i].real();
        continue;
      }
      std::vector<double> blk(2 * size_t(sz) * sz), ev(sz);
      for (int b = 0; b < sz; ++b)
        for (int a = 0; a < sz; ++a) {
          cd x = V[((ib + a) + size_t(nw) * (ib + b)) * 3 + i];
          cd y = V[((ib + b) + size_t(nw) * (ib + a)) * 3 + i];
          cd h = 0.5 * (x + std::conj(y));
          blk[2 * (a + size_t(sz) * b)] = h.real();
          blk[2 * (a + size_t(sz) * b) + 1] = h.imag();
        }
      if (orc_zheev(sz, blk.data(), ev.data()) < 0) return -1;
      for (int a = 0; a < sz; ++a) vg[3 * (ib + a) + i] = ev[a];
    }
    ib += sz;
  }
  return 0;
}

// Particle::getPopulation (particle.cpp:64-96).
double orc_population(int isBose, double energy, double temperature, double chemPot) {
  double population = 0.;
  if (temperature <= 0.) {
    if (energy <= chemPot) population = 1.;
    return population;
  }
  double y = (energy - chemPot) / temperature;
  if (isBose) {
    population = 1. / (std::exp(y) - 1);
    if (population < 0.) population = 0.;
  } else {
    population = 1. / (std::exp(y) + 1);
    if (population < 0.) population = 0.;
    if (population > 1.) population = 1.;
  }
  return population;
}

double orc_phonon_cutoff() { return 5. / kRyToCmm1; }  // el_scattering.cpp:127

// ---------------------------------------------------------------------------
// Full linewidth driver: ElScatteringMatrix::builder, switchCase 2
// (el_scattering.cpp:38-473) with inner == outer band structure on a
// Gamma-centred mesh with an energy window (storedAllQ3Case2,
// helper_el_scattering.cpp:58-162), no symmetries, no polar term.
// ---------------------------------------------------------------------------
struct orc_problem {
  double recip[9];  // reciprocal cell, column-major, columns b_i (Cartesian)
  int mesh[3];
  int nw, nRh;
  const double *Rh, *degh, *h0R;
  int nm, nRf;
  const double *Rf, *wf, *mat2R, *masses;
  int nRe, nRp;
  const double *Re, *degE, *Rp, *degP, *gW;
  int nCalc;
  const double *temperatures, *chemPots;
  double sigma, winMin, winMax;
  int natoms, nG;  // long-range term (nG = 0: none), same fields as phx_problem
  const double *atomPos, *bornCharges, *dielectric, *gVectors, *lrConstant;
  double lrNorm, gMax;
  int smearing;    // 0 Gaussian, 1 adaptive Gaussian
  // polar term of the coupling (0: off); the long-range fields' atoms, with
  // their own Born charges and dielectric tensor
  int polarOn;
  const double *polarPos, *polarBorn, *polarEps;
  double polarVolume;
  int polarCoarse[3];
};

struct orc_state {
  orc_problem p;
  int nk;                         // full mesh size
  std::vector<double> kcart;      // nk x 3
  std::vector<int> active;        // active -> full mesh index
  std::vector<int> bandLo, bandHi;
  std::vector<int> stateOffset;   // per active point, first iBte
  int numStates = 0;
  std::vector<double> energies;   // per state
  std::vector<double> stateV;     // per state: group velocity (adaptive smearing)
  std::vector<double> eigvecs;    // per active point: (nw, nb) col-major interleaved, packed
  std::vector<size_t> eigOffset;  // in complex elements
  std::vector<int> fullToActive;
  // phonons on the whole q mesh (index = full mesh index); only the q-set is filled
  std::vector<double> omega;      // nk x nm
  std::vector<double> phev;       // nk x (nm x nm) col-major interleaved
  std::vector<char> qInSet;
  std::vector<double> polarX;     // nk x nm complex: polarCorrectionPart1 of each q of the q-set
};

static void cart_of(const orc_problem& p, const double* x, double* c) {
  for (int i = 0; i < 3; ++i)
    c[i] = p.recip[i] * x[0] + p.recip[i + 3] * x[1] + p.recip[i + 6] * x[2];
}

// Solves recip * x = c for crystal coordinates x (Points::cartesianToCrystal,
// points.cpp:465-468).
static void crys_of(const orc_problem& p, const double* c, double* x) {
  const double* B = p.recip;
  double m[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m[i][j] = B[i + 3 * j];
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  double inv[3][3];
  inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
  inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
  inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
  inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
  for (int i = 0; i < 3; ++i) x[i] = inv[i][0] * c[0] + inv[i][1] * c[1] + inv[i][2] * c[2];
}

// Window::internalEnWindow (window.cpp:137-157) for one point: a band passes
// when winMin < e < winMax (both strict); the kept range is [first passing,
// last passing] (window.cpp:152-156), so bands in between are kept whether
// they pass or not.  Returns the number of passing bands; lo/hi are -1 when
// none passes.
int orc_window(int nb, const double* energies, double winMin, double winMax, int* lo, int* hi) {
  int count = 0;
  *lo = -1;
  *hi = -1;
  for (int ib = 0; ib < nb; ++ib) {
    double e = energies[ib];
    if (e < winMax && e > winMin) {
      if (*lo < 0) *lo = ib;
      *hi = ib;
      ++count;
    }
  }
  return count;
}

// Build band structure, window, q-set and phonons.  Returns a handle or null.
void* orc_setup(const orc_problem* prob) {
  auto* s = new orc_state();
  s->p = *prob;
  const orc_problem& p = s->p;
  s->nk = p.mesh[0] * p.mesh[1] * p.mesh[2];
  s->kcart.resize(3 * size_t(s->nk));
  std::vector<double> allE(size_t(s->nk) * p.nw), allU(2 * size_t(s->nk) * p.nw * p.nw);
  int fail = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : fail)
  for (int ik = 0; ik < s->nk; ++ik) {
    double x[3];
    orc_mesh_point(ik, p.mesh, x);
    cart_of(p, x, &s->kcart[3 * size_t(ik)]);
    fail += orc_el_diag(p.nw, p.nRh, p.Rh, p.degh, p.h0R, &s->kcart[3 * size_t(ik)],
                        &allE[size_t(ik) * p.nw], &allU[2 * size_t(ik) * p.nw * p.nw]) != 0;
  }
  if (fail) {
    delete s;
    return nullptr;
  }
  // Window::internalEnWindow (window.cpp:137-157) applied point by point in
  // ascending mesh order (active_bandstructure.cpp:826-890).
  s->fullToActive.assign(s->nk, -1);
  for (int ik = 0; ik < s->nk; ++ik) {
    int lo, hi;
    if (orc_window(p.nw, &allE[size_t(ik) * p.nw], p.winMin, p.winMax, &lo, &hi) == 0) continue;
    s->fullToActive[ik] = int(s->active.size());
    s->active.push_back(ik);
    s->bandLo.push_back(lo);
    s->bandHi.push_back(hi);
  }
  // states and sliced eigenvectors (active_bandstructure.cpp:981-1023)
  size_t eoff = 0;
  for (size_t a = 0; a < s->active.size(); ++a) {
    int ik = s->active[a];
    int nb = s->bandHi[a] - s->bandLo[a] + 1;
    s->stateOffset.push_back(s->numStates);
    s->numStates += nb;
    s->eigOffset.push_back(eoff);
    for (int ib = 0; ib < nb; ++ib) {
      int ibf = s->bandLo[a] + ib;
      s->energies.push_back(allE[size_t(ik) * p.nw + ibf]);
      for (int iw = 0; iw < p.nw; ++iw) {
        size_t src = 2 * ((size_t(ik) * p.nw * p.nw) + iw + size_t(p.nw) * ibf);
        s->eigvecs.push_back(allU[src]);
        s->eigvecs.push_back(allU[src + 1]);
      }
    }
    eoff += size_t(nb) * p.nw;
  }
  s->eigOffset.push_back(eoff);
  s->stateOffset.push_back(s->numStates);
  if (p.smearing == 1) {
    // group velocities of the window's bands (active_bandstructure.cpp:1032-1071)
    s->stateV.assign(3 * size_t(s->numStates), 0.);
    const int nact = int(s->active.size());
    int vfail = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : vfail)
    for (int a = 0; a < nact; ++a) {
      std::vector<double> vg(3 * size_t(p.nw));
      vfail += orc_el_group_velocities(p.nw, p.nRh, p.Rh, p.degh, p.h0R,
                                       &s->kcart[3 * size_t(s->active[a])], vg.data()) != 0;
      const int nb = s->bandHi[a] - s->bandLo[a] + 1;
      for (int b = 0; b < nb; ++b)
        for (int i = 0; i < 3; ++i)
          s->stateV[3 * size_t(s->stateOffset[a] + b) + i] = vg[3 * size_t(s->bandLo[a] + b) + i];
    }
    if (vfail) {
      delete s;
      return nullptr;
    }
  }
  // q-set: all folded k2 - k1 (helper_el_scattering.cpp:95-124)
  s->qInSet.assign(s->nk, 0);
  const int na = int(s->active.size());
  int offMesh = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : offMesh)
  for (int a2 = 0; a2 < na; ++a2)
    for (int a1 = 0; a1 < na; ++a1) {
      const double* k2 = &s->kcart[3 * size_t(s->active[a2])];
      const double* k1 = &s->kcart[3 * size_t(s->active[a1])];
      double qc[3] = {k2[0] - k1[0], k2[1] - k1[1], k2[2] - k1[2]}, qx[3];
      crys_of(p, qc, qx);
      int iq = orc_mesh_index(qx, p.mesh);
      if (iq < 0) {
        offMesh++;
        continue;
      }
#pragma omp atomic write
      s->qInSet[iq] = 1;
    }
  if (offMesh) {
    delete s;
    return nullptr;
  }
  s->omega.assign(size_t(s->nk) * p.nm, 0.);
  s->phev.assign(2 * size_t(s->nk) * p.nm * p.nm, 0.);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : fail)
  for (int iq = 0; iq < s->nk; ++iq) {
    if (!s->qInSet[iq]) continue;
    const orc_long_range lr{p.natoms, p.nG, p.atomPos, p.bornCharges, p.dielectric, p.gVectors,
                            p.lrConstant, p.lrNorm, p.gMax};
    fail += orc_ph_diag_lr(p.nm, p.nRf, p.Rf, p.wf, p.mat2R, p.masses, &s->kcart[3 * size_t(iq)],
                           p.nG > 0 ? &lr : nullptr, &s->omega[size_t(iq) * p.nm],
                           &s->phev[2 * size_t(iq) * p.nm * p.nm]) != 0;
  }
  if (fail) {
    delete s;
    return nullptr;
  }
  if (p.polarOn) {  // helper_el_scattering.cpp:170-195: Part1 at every q3 of the q-set
    s->polarX.assign(2 * size_t(s->nk) * p.nm, 0.);
    const int natoms = p.nm / 3;
#pragma omp parallel for schedule(dynamic, 16)
    for (int iq = 0; iq < s->nk; ++iq) {
      if (!s->qInSet[iq]) continue;
      orc_polar_part1(natoms, p.polarPos, p.polarBorn, p.polarEps, p.recip, p.polarVolume, p.polarCoarse, p.nm,
                      &s->kcart[3 * size_t(iq)], &s->phev[2 * size_t(iq) * p.nm * p.nm],
                      &s->polarX[2 * size_t(iq) * p.nm]);
    }
  }
  return s;
}

void orc_free(void* h) { delete static_cast<orc_state*>(h); }

int orc_num_active(void* h) { return int(static_cast<orc_state*>(h)->active.size()); }
int orc_num_states(void* h) { return static_cast<orc_state*>(h)->numStates; }

// active (full index), band extrema, state energies
void orc_get_window(void* h, int* active, int* lo, int* hi, double* energies) {
  auto* s = static_cast<orc_state*>(h);
  for (size_t a = 0; a < s->active.size(); ++a) {
    active[a] = s->active[a];
    lo[a] = s->bandLo[a];
    hi[a] = s->bandHi[a];
  }
  std::copy(s->energies.begin(), s->energies.end(), energies);
}

// Eigenvectors of one active point, (nw, nb) column-major.
void orc_get_eigvecs(void* h, int a, double* out) {
  auto* s = static_cast<orc_state*>(h);
  size_t n = s->eigOffset[a + 1] - s->eigOffset[a];
  std::copy(s->eigvecs.begin() + 2 * s->eigOffset[a],
            s->eigvecs.begin() + 2 * (s->eigOffset[a] + n), out);
}

// Phonon data on the q mesh (nk x nm and nk x nm x nm col-major).
void orc_get_phonons(void* h, double* omega, double* ev, char* inSet) {
  auto* s = static_cast<orc_state*>(h);
  std::copy(s->omega.begin(), s->omega.end(), omega);
  std::copy(s->phev.begin(), s->phev.end(), ev);
  std::copy(s->qInSet.begin(), s->qInSet.end(), inSet);
}

// Linewidths for the outer points listed in k1List (active indices; all if
// nk1 < 0), restricted to the inner points [k2Begin, k2End) of the active
// list (all if k2End < 0).  linewidth: (nCalc, numStates), accumulated into
// (caller zeroes it).  The loop body restates el_scattering.cpp:132-389.
// Returns the number of (k1,k2) pairs processed, or -1 on error.
long long orc_linewidth_pairs(void* h, int nk1, const int* k1List, int k2Begin, int k2End,
                              double* linewidth) {
  auto* s = static_cast<orc_state*>(h);
  const orc_problem& p = s->p;
  const int na = int(s->active.size());
  if (k2End < 0) k2End = na;
  std::vector<int> k1s;
  if (nk1 < 0) {
    k1s.resize(na);
    std::iota(k1s.begin(), k1s.end(), 0);
  } else {
    k1s.assign(k1List, k1List + nk1);
  }
  const double norm = 1. / (double(p.mesh[0]) * p.mesh[1] * p.mesh[2]);  // :64
  const double phononCutoff = orc_phonon_cutoff();
  // innerFermi (el_scattering.cpp:87-109)
  std::vector<double> fermi(size_t(p.nCalc) * s->numStates);
  for (int is = 0; is < s->numStates; ++is)
    for (int ic = 0; ic < p.nCalc; ++ic)
      fermi[size_t(ic) * s->numStates + is] =
          orc_population(0, s->energies[is], p.temperatures[ic], p.chemPots[ic]);
  std::vector<double> cached;
  long long pairs = 0;
  for (int a1 : k1s) {
    if (a1 < 0 || a1 >= na) return -1;
    const int nb1 = s->bandHi[a1] - s->bandLo[a1] + 1;
    const double* k1C = &s->kcart[3 * size_t(s->active[a1])];
    const double* U1 = &s->eigvecs[2 * s->eigOffset[a1]];
    const double* e1 = &s->energies[s->stateOffset[a1]];
    cached.assign(2 * size_t(p.nRp) * p.nm * nb1 * p.nw, 0.);
    orc_cache_elph(p.nw, p.nm, p.nRp, p.nRe, p.gW, p.Re, p.degE, U1, nb1, k1C, cached.data());
    const int nq = k2End - k2Begin;
    if (nq <= 0) continue;
    // preprocessing (el_scattering.cpp:194-210, helper_el_scattering.cpp:229-279)
    std::vector<int> nb2s(nq), iq3s(nq);
    std::vector<double> q3(3 * size_t(nq)), U2s, e3s(2 * size_t(nq) * p.nm * p.nm);
    std::vector<size_t> outoff(nq + 1, 0);
    for (int j = 0; j < nq; ++j) {
      int a2 = k2Begin + j;
      nb2s[j] = s->bandHi[a2] - s->bandLo[a2] + 1;
      const double* k2C = &s->kcart[3 * size_t(s->active[a2])];
      for (int i = 0; i < 3; ++i) q3[3 * j + i] = k2C[i] - k1C[i];
      double qx[3];
      crys_of(p, &q3[3 * j], qx);
      iq3s[j] = orc_mesh_index(qx, p.mesh);
      if (iq3s[j] < 0) return -1;
      U2s.insert(U2s.end(), s->eigvecs.begin() + 2 * s->eigOffset[a2],
                 s->eigvecs.begin() + 2 * s->eigOffset[a2 + 1]);
      std::copy(s->phev.begin() + 2 * size_t(iq3s[j]) * p.nm * p.nm,
                s->phev.begin() + 2 * size_t(iq3s[j] + 1) * p.nm * p.nm,
                e3s.begin() + 2 * size_t(j) * p.nm * p.nm);
      outoff[j + 1] = outoff[j] + size_t(nb1) * nb2s[j] * p.nm;
    }
    std::vector<double> coupling(outoff[nq]);
    std::vector<double> polar;
    if (p.polarOn) {
      polar.resize(2 * size_t(nq) * p.nm);
      for (int j = 0; j < nq; ++j)
        std::copy(s->polarX.begin() + 2 * size_t(iq3s[j]) * p.nm, s->polarX.begin() + 2 * size_t(iq3s[j] + 1) * p.nm,
                  polar.begin() + 2 * size_t(j) * p.nm);
    }
    orc_coupling_sq_polar(p.nw, p.nm, p.nRp, p.Rp, p.degP, cached.data(), nb1, nq, U2s.data(), nb2s.data(),
                          e3s.data(), q3.data(), U1, p.polarOn ? polar.data() : nullptr, coupling.data());
    // symmetrize coupling (el_scattering.cpp:216-224)
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < nq; ++j) {
      int a2 = k2Begin + j;
      orc_symmetrize_coupling(&coupling[outoff[j]], nb1, nb2s[j], p.nm, e1,
                              &s->energies[s->stateOffset[a2]], &s->omega[size_t(iq3s[j]) * p.nm]);
    }
    // postprocessing loop (el_scattering.cpp:228-386), serial
    std::vector<double> bose3(p.nCalc * size_t(p.nm));
    for (int j = 0; j < nq; ++j) {
      int a2 = k2Begin + j;
      const int nb2 = nb2s[j];
      const double* e2 = &s->energies[s->stateOffset[a2]];
      const double* w3 = &s->omega[size_t(iq3s[j]) * p.nm];
      for (int ic = 0; ic < p.nCalc; ++ic)
        for (int ib3 = 0; ib3 < p.nm; ++ib3)
          bose3[size_t(ic) * p.nm + ib3] = orc_population(1, w3[ib3], p.temperatures[ic], 0.);
      const double* cpl = &coupling[outoff[j]];
      for (int ib2 = 0; ib2 < nb2; ++ib2) {
        double en2 = e2[ib2];
        int iBte2 = s->stateOffset[a2] + ib2;
        for (int ib1 = 0; ib1 < nb1; ++ib1) {
          double en1 = e1[ib1];
          int iBte1 = s->stateOffset[a1] + ib1;
          for (int ib3 = 0; ib3 < p.nm; ++ib3) {
            double en3 = w3[ib3];
            if (en3 < phononCutoff) continue;
            double delta1, delta2;
            if (p.smearing == 1) {
              double v[3];
              for (int i = 0; i < 3; ++i)
                v[i] = s->stateV[3 * size_t(iBte1) + i] - s->stateV[3 * size_t(iBte2) + i];
              delta1 = orc_adaptive_delta(en1 - en2 + en3, v, p.recip, p.mesh);
              delta2 = orc_adaptive_delta(en1 - en2 - en3, v, p.recip, p.mesh);
            } else {
              delta1 = orc_gaussian_delta(en1 - en2 + en3, p.sigma);
              delta2 = orc_gaussian_delta(en1 - en2 - en3, p.sigma);
            }
            if (delta1 <= 0. && delta2 <= 0.) continue;
            double c = cpl[ib1 + size_t(nb1) * (ib2 + size_t(nb2) * ib3)];
            for (int ic = 0; ic < p.nCalc; ++ic) {
              double fermi2 = fermi[size_t(ic) * s->numStates + iBte2];
              double b3 = bose3[size_t(ic) * p.nm + ib3];
              double rate = c * ((fermi2 + b3) * delta1 + (1. - fermi2 + b3) * delta2) *
                            norm / en3 * kPi;
              linewidth[size_t(ic) * s->numStates + iBte1] += rate;
            }
          }
        }
      }
    }
    pairs += nq;
  }
  return pairs;
}

// ScatteringMatrix::degeneracyAveragingLinewidths (scattering.cpp:1168-1207)
// at one point: each band takes the mean of the bands within 1e-6 Ry of its
// own energy (not a transitive group), all means computed from the values
// before averaging, divided by float(n).  lw points at the point's first state
// of row iCalc = 0; other rows are not touched (scattering.cpp:1192,1204).
void orc_average_point(int nb, const double* energies, double* lw) {
  std::vector<double> tmp(nb, 0.);
  for (int ib1 = 0; ib1 < nb; ++ib1) {
    double ekk = energies[ib1];
    int n = 0;
    double t2 = 0.;
    for (int ib2 = 0; ib2 < nb; ++ib2) {
      if (std::abs(energies[ib2] - ekk) < 1.0e-6) {
        n++;
        t2 = t2 + lw[ib2];
      }
    }
    tmp[ib1] = t2 / float(n);
  }
  for (int ib1 = 0; ib1 < nb; ++ib1) lw[ib1] = tmp[ib1];
}

// Every active point of a set-up problem; linewidth is (nCalc, numStates).
void orc_degeneracy_average(void* h, double* linewidth) {
  auto* s = static_cast<orc_state*>(h);
  for (size_t a = 0; a < s->active.size(); ++a) {
    int off = s->stateOffset[a];
    orc_average_point(s->stateOffset[a + 1] - off, &s->energies[off], linewidth + off);
  }
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}


// ---------------------------------------------------------------------------
// PhElScatteringMatrix::builder (phel_scattering.cpp:43-363): phonon
// linewidths on the Gamma-centred q mesh qmesh from the electrons of the k mesh
// (window as the electron builder's), k2 = k1 + q3 with all bands, coupling
// symmetrized, rate |g|^2 f(1-f)(e1) delta(e1 - e2 + w3) norm / T pi with
// norm = 1 / prod(kMesh), then degeneracy averaging over the phonon modes
// (row iCalc = 0, scattering.cpp:1168-1207).  out: nCalc x nq x nm.
// Returns the number of (k1, q3) pairs, or -1.
long long orc_phel(const orc_problem* prob, const int* qmesh, int nk1, const int* k1_list, int average,
                   double* out, double* loop_seconds) {
  const orc_problem& p = *prob;
  const int nw = p.nw, nm = p.nm;
  const int nk = p.mesh[0] * p.mesh[1] * p.mesh[2];
  const int nq = qmesh[0] * qmesh[1] * qmesh[2];
  const double norm = 1. / double(nk);
  const double phononCutoff = orc_phonon_cutoff();
  // electrons on the k mesh, window
  std::vector<double> kcart(3 * size_t(nk)), allE(size_t(nk) * nw), allU(2 * size_t(nk) * nw * nw);
  int fail = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : fail)
  for (int ik = 0; ik < nk; ++ik) {
    double x[3];
    orc_mesh_point(ik, p.mesh, x);
    cart_of(p, x, &kcart[3 * size_t(ik)]);
    fail += orc_el_diag(nw, p.nRh, p.Rh, p.degh, p.h0R, &kcart[3 * size_t(ik)], &allE[size_t(ik) * nw],
                        &allU[2 * size_t(ik) * nw * nw]) != 0;
  }
  if (fail) return -1;
  std::vector<int> active, lo, hi;
  for (int ik = 0; ik < nk; ++ik) {
    int l, h;
    if (orc_window(nw, &allE[size_t(ik) * nw], p.winMin, p.winMax, &l, &h) == 0) continue;
    active.push_back(ik);
    lo.push_back(l);
    hi.push_back(h);
  }
  // phonons on the q mesh
  std::vector<double> qcart(3 * size_t(nq)), w(size_t(nq) * nm), ev(2 * size_t(nq) * nm * nm);
  const orc_long_range lr{p.natoms, p.nG, p.atomPos, p.bornCharges, p.dielectric, p.gVectors,
                          p.lrConstant, p.lrNorm, p.gMax};
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : fail)
  for (int iq = 0; iq < nq; ++iq) {
    double x[3];
    orc_mesh_point(iq, qmesh, x);
    cart_of(p, x, &qcart[3 * size_t(iq)]);
    fail += orc_ph_diag_lr(nm, p.nRf, p.Rf, p.wf, p.mat2R, p.masses, &qcart[3 * size_t(iq)],
                           p.nG > 0 ? &lr : nullptr, &w[size_t(iq) * nm], &ev[2 * size_t(iq) * nm * nm]) != 0;
  }
  if (fail) return -1;
  std::vector<double> lw(size_t(p.nCalc) * nq * nm, 0.);
  long long pairs = 0;
  std::vector<double> cached(2 * size_t(p.nRp) * nm * nw * nw);
  // outer points: k1_list (active indices) or all of them
  std::vector<int> list;
  if (nk1 < 0) {
    for (size_t a = 0; a < active.size(); ++a) list.push_back(int(a));
  } else {
    for (int i = 0; i < nk1; ++i) {
      if (k1_list[i] < 0 || size_t(k1_list[i]) >= active.size()) return -1;
      list.push_back(k1_list[i]);
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (int a : list) {
    const int ik1 = active[a], nb1 = hi[a] - lo[a] + 1;
    const double* k1 = &kcart[3 * size_t(ik1)];
    std::vector<double> e1(nb1), U1(2 * size_t(nw) * nb1), v1(3 * size_t(nb1), 0.);
    for (int b = 0; b < nb1; ++b) {
      e1[b] = allE[size_t(ik1) * nw + lo[a] + b];
      for (int iw = 0; iw < nw; ++iw) {
        const size_t src = 2 * (size_t(ik1) * nw * nw + iw + size_t(nw) * (lo[a] + b));
        U1[2 * (iw + size_t(nw) * b)] = allU[src];
        U1[2 * (iw + size_t(nw) * b) + 1] = allU[src + 1];
      }
    }
    if (p.smearing == 1) {
      std::vector<double> vg(3 * size_t(nw));
      if (orc_el_group_velocities_rule(nw, p.nRh, p.Rh, p.degh, p.h0R, k1, 1, vg.data()) != 0) return -1;
      for (int b = 0; b < nb1; ++b)
        for (int i = 0; i < 3; ++i) v1[3 * b + i] = vg[3 * (lo[a] + b) + i];
    }
    orc_cache_elph(nw, nm, p.nRp, p.nRe, p.gW, p.Re, p.degE, U1.data(), nb1, k1, cached.data());
    std::vector<double> E2(size_t(nq) * nw), U2(2 * size_t(nq) * nw * nw), V2;
    std::vector<int> nb2s(nq, nw);
    int f2 = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : f2)
    for (int q = 0; q < nq; ++q) {
      double k2[3];
      for (int i = 0; i < 3; ++i) k2[i] = qcart[3 * size_t(q) + i] + k1[i];
      f2 += orc_el_diag(nw, p.nRh, p.Rh, p.degh, p.h0R, k2, &E2[size_t(q) * nw],
                        &U2[2 * size_t(q) * nw * nw]) != 0;
    }
    if (f2) return -1;
    if (p.smearing == 1) {
      V2.assign(3 * size_t(nq) * nw, 0.);
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : f2)
      for (int q = 0; q < nq; ++q) {
        double k2[3];
        for (int i = 0; i < 3; ++i) k2[i] = qcart[3 * size_t(q) + i] + k1[i];
        f2 += orc_el_group_velocities_rule(nw, p.nRh, p.Rh, p.degh, p.h0R, k2, 0, &V2[3 * size_t(q) * nw]) != 0;
      }
      if (f2) return -1;
    }
    std::vector<double> coupling(size_t(nq) * nb1 * nw * nm);
    orc_coupling_sq(nw, nm, p.nRp, p.Rp, p.degP, cached.data(), nb1, nq, U2.data(), nb2s.data(),
                    ev.data(), qcart.data(), coupling.data());
    for (int q = 0; q < nq; ++q) {
      double* c = &coupling[size_t(q) * nb1 * nw * nm];
      const double* e2 = &E2[size_t(q) * nw];
      const double* w3 = &w[size_t(q) * nm];
      orc_symmetrize_coupling(c, nb1, nw, nm, e1.data(), e2, w3);
      for (int ib3 = 0; ib3 < nm; ++ib3)
        for (int ic = 0; ic < p.nCalc; ++ic) {
          for (int ib1 = 0; ib1 < nb1; ++ib1)
            for (int ib2 = 0; ib2 < nw; ++ib2) {
              const double en3 = w3[ib3];
              double delta = 0.;
              if (en3 >= phononCutoff) {
                const double en = e1[ib1] - e2[ib2] + en3;
                if (p.smearing == 1) {
                  double v[3];
                  for (int i = 0; i < 3; ++i) v[i] = v1[3 * ib1 + i] - V2[(size_t(q) * nw + ib2) * 3 + i];
                  delta = orc_adaptive_delta(en, v, p.recip, p.mesh);
                } else {
                  delta = orc_gaussian_delta(en, p.sigma);
                }
                delta = std::max(delta, 0.);
              }
              // Particle::getPopPopPm1 (particle.cpp:117-128)
              const double T = p.temperatures[ic];
              const double x = std::cosh((e1[ib1] - p.chemPots[ic]) / 2. / T);
              const double F = 0.25 / x / x;
              const double rate = c[ib1 + size_t(nb1) * (ib2 + size_t(nw) * ib3)] * F * delta * norm / T * kPi;
              lw[(size_t(ic) * nq + q) * nm + ib3] += rate;
            }
        }
    }
    pairs += nq;
  }
  if (loop_seconds)
    *loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (average)  // the phonon modes of each q3, row iCalc = 0
    for (int q = 0; q < nq; ++q) orc_average_point(nm, &w[size_t(q) * nm], &lw[size_t(q) * nm]);
  std::copy(lw.begin(), lw.end(), out);
  return pairs;
}

}  // extern "C"
