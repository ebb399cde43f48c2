// Band velocities (velocity operator in the band basis) for electrons and
// phonons on gfx950.
//
//  * electrons: ElectronH0Wannier::kokkosBatchedDiagonalizeWithVelocities
//    (electron_h0_wannier.cpp:888-999): der_i = (H(k + d e_i) - H(k - d e_i)) / 2d,
//    hermitized, v_i = U^H der_i U;
//  * phonons: PhononH0::kokkosBatchedDiagonalizeWithVelocities
//    (phonon_h0_kokkos.cpp:378-532): the same with sqrt(D) = X diag(omega) X^H
//    rebuilt from the eigen-decompositions at q +- d e_i (X: eigenvectors
//    without the mass scaling);
//  * both: zero at Gamma, then every degenerate block
//    (energies within a threshold of the block's first band) is diagonalized
//    per Cartesian direction and replaced by V^H B V
//    (HarmonicHamiltonian::kokkosBatchedTreatDegenerateVelocities,
//    harmonic.cpp:3-203).
//
// "Gamma" has two definitions in the reference: the batched GPU forms zero the
// velocities unless |k|^2 > 1e-6 (electron_h0_wannier.cpp:978-983,
// phonon_h0_kokkos.cpp:499-506); the single-point CPU forms only when
// |k| < 1e-6 (electron_h0_wannier.cpp:682-686, phonon_h0.cpp:1057-1060).
// gammaRule 0 / 1 selects them.
//
// k_vel_operator: one workgroup per (point, direction); U, der and the
// intermediate product live in LDS (3 n^2 complex, 93 KB at n = 44).
// k_vel_degenerate: one thread per (point, direction), complex Jacobi on the
// block in registers (blocks of at most kVelDegMax bands).
#include "phx_common.hpp"

namespace phx {

constexpr int kVelNMax = 44;
constexpr int kVelDegMax = 8;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconjmul(double2 a, double2 b) {  // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ inline bool vel_gamma(const double* kc, int p, int gammaRule) {
  const double k2 = kc[3 * p] * kc[3 * p] + kc[3 * p + 1] * kc[3 * p + 1] + kc[3 * p + 2] * kc[3 * p + 2];
  return gammaRule ? (k2 < 1.0e-12) : !(k2 > 1.0e-6);
}

// H7 / X7: (7 np) matrices n x n column-major; point p's central matrix is 7p,
// direction i's +d / -d matrices are 7p + 2i + 1 / 7p + 2i + 2.
template <bool PHONON>
__global__ void __launch_bounds__(256) k_vel_operator(
    int np, int n, double delta, int gammaRule, const double* __restrict__ kc,  // np x 3 (central points)
    const double2* __restrict__ Uc,     // electrons: np central eigenvector matrices
    const double2* __restrict__ M7,     // electrons: H at 7 np points; phonons: e / sqrt(M) at 7 np points
    const double* __restrict__ W7,      // phonons: omega at 7 np points
    const double* __restrict__ masses,  // phonons: n
    double2* __restrict__ vel) {        // np x n x n x 3, (p, m, n, i) row-major
  extern __shared__ double2 lds[];
  double2* Us = lds;
  double2* Ds = lds + n * n;
  double2* Ts = lds + 2 * n * n;
  const int p = blockIdx.x, dir = blockIdx.y;
  const int nn = n * n;
  const size_t c = PHONON ? size_t(7) * p : size_t(p);
  const double2* U = PHONON ? M7 + c * nn : Uc + c * nn;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    double2 u = U[e];
    if (PHONON) {
      const double s = sqrt(masses[e % n]);
      u.x *= s;
      u.y *= s;
    }
    Us[e] = u;
  }
  const size_t ip = size_t(7) * p + 2 * dir + 1, im = ip + 1;
  if (!PHONON) {
    const double2* Hp = M7 + ip * nn;
    const double2* Hm = M7 + im * nn;
    const double f = 0.25 / delta;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
      const int r = e % n, col = e / n;
      const double2 a = Hp[e], b = Hm[e];
      const double2 at = Hp[col + size_t(n) * r], bt = Hm[col + size_t(n) * r];
      Ds[e] = make_double2(f * ((a.x - b.x) + (at.x - bt.x)), f * ((a.y - b.y) - (at.y - bt.y)));
    }
  } else {
    const double2* Xp = M7 + ip * nn;
    const double2* Xm = M7 + im * nn;
    const double* wp = W7 + ip * n;
    const double* wm = W7 + im * n;
    const double f = 1. / (2. * delta);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
      const int r = e % n, col = e / n;
      const double sr = sqrt(masses[r]), sc = sqrt(masses[col]);
      double2 acc = make_double2(0., 0.);
      for (int l = 0; l < n; ++l) {
        double2 xr = Xp[r + size_t(n) * l], xc = Xp[col + size_t(n) * l];
        double2 t = cmul(make_double2(xr.x * sr, xr.y * sr), make_double2(xc.x * sc, -xc.y * sc));
        acc.x += wp[l] * t.x;
        acc.y += wp[l] * t.y;
        xr = Xm[r + size_t(n) * l];
        xc = Xm[col + size_t(n) * l];
        t = cmul(make_double2(xr.x * sr, xr.y * sr), make_double2(xc.x * sc, -xc.y * sc));
        acc.x -= wm[l] * t.x;
        acc.y -= wm[l] * t.y;
      }
      Ts[e] = make_double2(f * acc.x, f * acc.y);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
      const int r = e % n, col = e / n;
      const double2 a = Ts[e], b = Ts[col + n * r];
      Ds[e] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    }
  }
  __syncthreads();
  // Ts = U^H Ds
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const int r = e % n, col = e / n;
    double2 acc = make_double2(0., 0.);
    for (int l = 0; l < n; ++l) {
      const double2 t = cconjmul(Us[l + n * r], Ds[l + n * col]);
      acc.x += t.x;
      acc.y += t.y;
    }
    Ts[e] = acc;
  }
  __syncthreads();
  const bool gamma = vel_gamma(kc, p, gammaRule);
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const int r = e % n, col = e / n;
    double2 acc = make_double2(0., 0.);
    if (!gamma) {
      for (int l = 0; l < n; ++l) {
        const double2 t = cmul(Ts[r + n * l], Us[l + n * col]);
        acc.x += t.x;
        acc.y += t.y;
      }
    }
    vel[((size_t(p) * n + r) * n + col) * 3 + dir] = acc;
  }
}

// Diagonalize each degenerate block of v_i (hermitized) and write back
// V^H B V with the eigenvalues ascending on its diagonal.
__global__ void __launch_bounds__(64) k_vel_degenerate(int np, int n, double threshold, int gammaRule,
                                                       const double* __restrict__ E,  // np x n (ascending)
                                                       const double* __restrict__ kc,
                                                       double2* __restrict__ vel, int* __restrict__ tooBig) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * np) return;
  const int p = t / 3, dir = t % 3;
  if (vel_gamma(kc, p, gammaRule)) return;  // velocities are zero at Gamma
  const double* e = E + size_t(p) * n;
  auto V = [&](int a, int b) -> double2& { return vel[((size_t(p) * n + a) * n + b) * 3 + dir]; };
  for (int ib = 0; ib < n;) {
    int s = 1;
    while (ib + s < n && fabs(e[ib] - e[ib + s]) <= threshold) ++s;
    if (s > 1) {
      if (s > kVelDegMax) {
        atomicMax(tooBig, s);
        ib += s;
        continue;
      }
      double2 A[kVelDegMax][kVelDegMax];
      double fro = 0.;
      for (int a = 0; a < s; ++a)
        for (int b = 0; b < s; ++b) {
          const double2 x = V(ib + a, ib + b), y = V(ib + b, ib + a);
          A[a][b] = make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
          fro += A[a][b].x * A[a][b].x + A[a][b].y * A[a][b].y;
        }
      for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.;
        for (int a = 0; a < s; ++a)
          for (int b = a + 1; b < s; ++b) off += A[a][b].x * A[a][b].x + A[a][b].y * A[a][b].y;
        if (off <= 1e-34 * fro || off == 0.) break;
        for (int a = 0; a < s - 1; ++a)
          for (int b = a + 1; b < s; ++b) {
            const double r = sqrt(A[a][b].x * A[a][b].x + A[a][b].y * A[a][b].y);
            if (r == 0.) continue;
            // J = diag(1, e^{-i phi}) on (a, b) followed by the real rotation of the 2x2 block
            const double cph = A[a][b].x / r, sph = A[a][b].y / r;
            const double tau = (A[b][b].x - A[a][a].x) / (2. * r);
            const double tt = (tau >= 0. ? 1. : -1.) / (fabs(tau) + sqrt(1. + tau * tau));
            const double cs = 1. / sqrt(1. + tt * tt), sn = tt * cs;
            const double2 jaa = make_double2(cs, 0.), jab = make_double2(sn, 0.);
            const double2 jba = make_double2(-sn * cph, sn * sph);   // -s e^{-i phi}
            const double2 jbb = make_double2(cs * cph, -cs * sph);   //  c e^{-i phi}
            for (int k = 0; k < s; ++k) {  // columns: A <- A J
              const double2 xa = A[k][a], xb = A[k][b];
              const double2 na = cmul(xa, jaa), nb2 = cmul(xb, jba);
              const double2 ma = cmul(xa, jab), mb = cmul(xb, jbb);
              A[k][a] = make_double2(na.x + nb2.x, na.y + nb2.y);
              A[k][b] = make_double2(ma.x + mb.x, ma.y + mb.y);
            }
            for (int k = 0; k < s; ++k) {  // rows: A <- J^H A
              const double2 xa = A[a][k], xb = A[b][k];
              const double2 na = cconjmul(jaa, xa), nb2 = cconjmul(jba, xb);
              const double2 ma = cconjmul(jab, xa), mb = cconjmul(jbb, xb);
              A[a][k] = make_double2(na.x + nb2.x, na.y + nb2.y);
              A[b][k] = make_double2(ma.x + mb.x, ma.y + mb.y);
            }
            A[a][b] = make_double2(0., 0.);
            A[b][a] = make_double2(0., 0.);
          }
      }
      // ascending order of the eigenvalues
      int perm[kVelDegMax];
      for (int a = 0; a < s; ++a) perm[a] = a;
      for (int a = 1; a < s; ++a) {
        const int x = perm[a];
        int b = a - 1;
        while (b >= 0 && A[perm[b]][perm[b]].x > A[x][x].x) {
          perm[b + 1] = perm[b];
          --b;
        }
        perm[b + 1] = x;
      }
      for (int a = 0; a < s; ++a)
        for (int b = 0; b < s; ++b) V(ib + a, ib + b) = A[perm[a]][perm[b]];
    }
    ib += s;
  }
}

void launch_vel_operator(bool phonon, int np, int n, double delta, int gammaRule, const double* kc,
                         const double2* Uc,
                         const double2* M7, const double* W7, const double* masses, double2* vel,
                         hipStream_t s) {
  if (np <= 0) return;
  if (n < 1 || n > kVelNMax) throw Error(-1, "velocities: matrix size must be 1..44");
  const size_t lds = size_t(3) * n * n * sizeof(double2);
  dim3 grid(np, 3);
  auto kern = phonon ? k_vel_operator<true> : k_vel_operator<false>;
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, np, n, delta, gammaRule, kc, Uc, M7, W7, masses, vel);
  PHX_HIP(hipGetLastError());
}

void launch_vel_degenerate(int np, int n, double threshold, int gammaRule, const double* E,
                           const double* kc, double2* vel, int* tooBig, hipStream_t s) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_vel_degenerate, dim3(ceil_div(3LL * np, 64)), dim3(64), 0, s, np, n, threshold,
                     gammaRule, E, kc, vel, tooBig);
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
