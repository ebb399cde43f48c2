// Batched Hermitian eigensolver: one matrix per wavefront, LDS resident.
//
// Replaces kokkosZHEEV (common_kokkos.cpp:11-100): eigenvalues ascending,
// eigenvectors written over the input as the columns of a column-major
// matrix (each matrix contiguous, LayoutStride (N^2,1,N), common_kokkos.cpp:
// 32-41; electron_h0_wannier.cpp:787-791).  The algorithm is a parallel
// cyclic complex Jacobi method with round-robin pair ordering: all N/2
// disjoint rotations of a round are applied at once by the 64 lanes, the
// matrix and the accumulated eigenvectors stay in LDS for all sweeps.
//
// The phonon variant fuses the steps either side of the eigensolve in
// PhononH0::kokkosBatchedBuildBlochHamiltonian / DiagonalizeFromCoordinates
// (phonon_h0_kokkos.cpp:150-170, 186-219): D <- (D + D^H)/2 / sqrt(M_m M_n)
// before, omega = sign(l) sqrt|l| and e_m <- e_m / sqrt(M_m) after.
#include "phx_common.hpp"

namespace phx {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// One rotation of a round: J = [[c, s e^{i phi}], [-s e^{-i phi}, c]] on (p, q).
struct Rot {
  int p, q;
  double c;
  double2 sp;  // s e^{+i phi}
  double2 sm;  // s e^{-i phi}
};

__device__ inline int rr_player(int pos, int round, int n) {
  // circle method: player 0 fixed, the others rotate by one per round
  if (pos == 0) return 0;
  return (round + pos - 1) % (n - 1) + 1;
}

template <bool PHONON>
__global__ void __launch_bounds__(64) k_jacobi(int N, int M, double2* __restrict__ A,
                                               double* __restrict__ W, int* __restrict__ info,
                                               const double* __restrict__ masses, int maxSweeps) {
  extern __shared__ double2 lds[];
  const int mat = blockIdx.x;
  if (mat >= M) return;
  const int lane = threadIdx.x;
  const int NN = N * N;
  double2* a = lds;
  double2* v = lds + NN;
  Rot* rot = reinterpret_cast<Rot*>(v + NN);
  double* red = reinterpret_cast<double*>(rot + 32);
  double2* Am = A + size_t(mat) * NN;

  // load (column-major) with hermitization
  for (int e = lane; e < NN; e += 64) {
    int i = e % N, j = e / N;
    double2 x;
    if (PHONON) {
      double2 aij = Am[i + N * j], aji = Am[j + N * i];
      x = make_double2(0.5 * (aij.x + aji.x), 0.5 * (aij.y - aji.y));
      x.x /= sqrt(masses[i] * masses[j]);
      x.y /= sqrt(masses[i] * masses[j]);
    } else {
      // lower triangle is the reference (uplo = L)
      x = (i >= j) ? Am[i + N * j] : cconj(Am[j + N * i]);
      if (i == j) x.y = 0.;
    }
    a[e] = x;
    v[e] = make_double2(i == j ? 1. : 0., 0.);
  }
  __syncthreads();

  double part = 0.;
  for (int e = lane; e < NN; e += 64) part += a[e].x * a[e].x + a[e].y * a[e].y;
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
  const double norm2 = part;

  const int n = N + (N & 1);  // even number of players (N is a dummy when odd)
  const int npairs = n / 2;
  int sweep = 0;
  bool converged = false;
  for (; sweep < maxSweeps; ++sweep) {
    double off = 0.;
    for (int e = lane; e < NN; e += 64) {
      int i = e % N, j = e / N;
      if (i != j) off += a[e].x * a[e].x + a[e].y * a[e].y;
    }
    for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
    if (off <= 1e-30 * norm2) {
      converged = true;
      break;
    }
    for (int r = 0; r < n - 1; ++r) {
      if (lane < npairs) {
        int p = rr_player(lane, r, n), q = rr_player(n - 1 - lane, r, n);
        if (p > q) { int t = p; p = q; q = t; }
        Rot R;
        R.p = p;
        R.q = q;
        R.c = 1.;
        R.sp = make_double2(0., 0.);
        R.sm = make_double2(0., 0.);
        if (q < N) {
          double2 apq = a[p + N * q];
          double rr = sqrt(apq.x * apq.x + apq.y * apq.y);
          if (rr > 1e-300) {
            double app = a[p + N * p].x, aqq = a[q + N * q].x;
            double tau = (aqq - app) / (2. * rr);
            double t = (tau >= 0. ? 1. : -1.) / (fabs(tau) + sqrt(1. + tau * tau));
            double c = 1. / sqrt(1. + t * t);
            double s = t * c;
            double2 ph = make_double2(apq.x / rr, apq.y / rr);
            R.c = c;
            R.sp = make_double2(s * ph.x, s * ph.y);
            R.sm = make_double2(s * ph.x, -s * ph.y);
          } else {
            R.q = N;  // nothing to do
          }
        }
        rot[lane] = R;
      }
      __syncthreads();
      // A <- A J and V <- V J (columns p, q of every pair)
      for (int it = lane; it < N * npairs; it += 64) {
        int k = it % N, pr = it / N;
        Rot R = rot[pr];
        if (R.q >= N) continue;
        double2 akp = a[k + N * R.p], akq = a[k + N * R.q];
        double2 t1 = cmul(akq, R.sm), t2 = cmul(akp, R.sp);
        a[k + N * R.p] = make_double2(akp.x * R.c - t1.x, akp.y * R.c - t1.y);
        a[k + N * R.q] = make_double2(t2.x + akq.x * R.c, t2.y + akq.y * R.c);
        double2 vkp = v[k + N * R.p], vkq = v[k + N * R.q];
        t1 = cmul(vkq, R.sm);
        t2 = cmul(vkp, R.sp);
        v[k + N * R.p] = make_double2(vkp.x * R.c - t1.x, vkp.y * R.c - t1.y);
        v[k + N * R.q] = make_double2(t2.x + vkq.x * R.c, t2.y + vkq.y * R.c);
      }
      __syncthreads();
      // A <- J^H A (rows p, q of every pair)
      for (int it = lane; it < N * npairs; it += 64) {
        int k = it % N, pr = it / N;
        Rot R = rot[pr];
        if (R.q >= N) continue;
        double2 apk = a[R.p + N * k], aqk = a[R.q + N * k];
        double2 t1 = cmul(R.sp, aqk), t2 = cmul(R.sm, apk);
        a[R.p + N * k] = make_double2(R.c * apk.x - t1.x, R.c * apk.y - t1.y);
        a[R.q + N * k] = make_double2(t2.x + R.c * aqk.x, t2.y + R.c * aqk.y);
      }
      __syncthreads();
      if (lane < npairs) {
        Rot R = rot[lane];
        if (R.q < N) {
          a[R.p + N * R.q] = make_double2(0., 0.);
          a[R.q + N * R.p] = make_double2(0., 0.);
          a[R.p + N * R.p].y = 0.;
          a[R.q + N * R.q].y = 0.;
        }
      }
      __syncthreads();
    }
  }
  // ascending order by rank (stable)
  if (lane < N) red[lane] = a[lane + N * lane].x;
  __syncthreads();
  for (int i = lane; i < N; i += 64) {
    double di = red[i];
    int rank = 0;
    for (int j = 0; j < N; ++j) {
      double dj = red[j];
      rank += (dj < di) || (dj == di && j < i);
    }
    double lam = di;
    if (PHONON) lam = lam > 0. ? sqrt(lam) : -sqrt(-lam);
    W[size_t(mat) * N + rank] = lam;
    for (int k = 0; k < N; ++k) {
      double2 x = v[k + N * i];
      if (PHONON) {
        double s = sqrt(masses[k]);
        x.x /= s;
        x.y /= s;
      }
      Am[k + N * rank] = x;
    }
  }
  if (lane == 0 && info) info[mat] = converged ? sweep : -1;
}

size_t jacobi_lds_bytes(int N) {
  return size_t(2) * N * N * sizeof(double2) + 32 * sizeof(Rot) + 64 * sizeof(double);
}

// The same Jacobi method with the order N fixed at compile time and 64 / L
// matrices per wavefront, L >= N lanes each: every lane works on one matrix,
// index arithmetic is by constants, and a wave issues its instructions for
// several matrices at once.  Each matrix goes through the rotations of the
// one-matrix kernel, in the same order, with the same arithmetic per
// element; the wave runs sweeps until all its matrices have converged (a
// converged matrix's rotations are skipped, so its result and sweep count
// are its own).
template <bool PHONON, int N, int L>
__global__ void __launch_bounds__(64) k_jacobi_n(int M, double2* __restrict__ A, double* __restrict__ W,
                                                 int* __restrict__ info, const double* __restrict__ masses,
                                                 int maxSweeps) {
  constexpr int G = 64 / L, NN = N * N, NE = N + (N & 1), NP = NE / 2;
  static_assert(L >= N && 64 % L == 0, "a matrix needs at least N lanes");
  extern __shared__ double2 lds[];
  const int lane = threadIdx.x, sg = lane / L, li = lane % L;
  const int mat = blockIdx.x * G + sg;
  const bool on = mat < M;
  double2* a = lds + size_t(sg) * 2 * NN;
  double2* v = a + NN;
  Rot* rot = reinterpret_cast<Rot*>(lds + size_t(G) * 2 * NN) + sg * NP;
  double* red = reinterpret_cast<double*>(reinterpret_cast<Rot*>(lds + size_t(G) * 2 * NN) + G * NP) + sg * N;
  double2* Am = A + size_t(on ? mat : 0) * NN;

  for (int e = li; e < NN; e += L) {
    const int i = e % N, j = e / N;
    double2 x = make_double2(0., 0.);
    if (on) {
      if (PHONON) {
        const double2 aij = Am[i + N * j], aji = Am[j + N * i];
        x = make_double2(0.5 * (aij.x + aji.x), 0.5 * (aij.y - aji.y));
        x.x /= sqrt(masses[i] * masses[j]);
        x.y /= sqrt(masses[i] * masses[j]);
      } else {
        x = (i >= j) ? Am[i + N * j] : cconj(Am[j + N * i]);
        if (i == j) x.y = 0.;
      }
    }
    a[e] = x;
    v[e] = make_double2(i == j ? 1. : 0., 0.);
  }
  __syncthreads();
  double part = 0.;
  for (int e = li; e < NN; e += L) part += a[e].x * a[e].x + a[e].y * a[e].y;
  for (int o = L / 2; o > 0; o >>= 1) part += __shfl_xor(part, o);
  const double norm2 = part;

  int sweep = 0, mySweeps = -1;
  for (; sweep < maxSweeps; ++sweep) {
    double off = 0.;
    for (int e = li; e < NN; e += L) {
      const int i = e % N, j = e / N;
      if (i != j) off += a[e].x * a[e].x + a[e].y * a[e].y;
    }
    for (int o = L / 2; o > 0; o >>= 1) off += __shfl_xor(off, o);
    if (mySweeps < 0 && off <= 1e-30 * norm2) mySweeps = sweep;
    if (__all(mySweeps >= 0)) break;
    const bool active = mySweeps < 0;
    for (int r = 0; r < NE - 1; ++r) {
      if (li < NP) {
        int pp = rr_player(li, r, NE), qq = rr_player(NE - 1 - li, r, NE);
        if (pp > qq) { const int t = pp; pp = qq; qq = t; }
        Rot R;
        R.p = pp;
        R.q = qq;
        R.c = 1.;
        R.sp = make_double2(0., 0.);
        R.sm = make_double2(0., 0.);
        if (qq < N && active) {
          const double2 apq = a[pp + N * qq];
          const double rr = sqrt(apq.x * apq.x + apq.y * apq.y);
          if (rr > 1e-300) {
            const double app = a[pp + N * pp].x, aqq = a[qq + N * qq].x;
            const double tau = (aqq - app) / (2. * rr);
            const double t = (tau >= 0. ? 1. : -1.) / (fabs(tau) + sqrt(1. + tau * tau));
            const double c = 1. / sqrt(1. + t * t);
            const double sn = t * c;
            const double2 ph = make_double2(apq.x / rr, apq.y / rr);
            R.c = c;
            R.sp = make_double2(sn * ph.x, sn * ph.y);
            R.sm = make_double2(sn * ph.x, -sn * ph.y);
          } else {
            R.q = N;
          }
        } else {
          R.q = N;  // dummy player, or a converged matrix
        }
        rot[li] = R;
      }
      __syncthreads();
      // A <- A J and V <- V J (columns p, q of every pair)
      for (int it = li; it < N * NP; it += L) {
        const int k = it % N, pr = it / N;
        const Rot R = rot[pr];
        if (R.q >= N) continue;
        const double2 akp = a[k + N * R.p], akq = a[k + N * R.q];
        double2 t1 = cmul(akq, R.sm), t2 = cmul(akp, R.sp);
        a[k + N * R.p] = make_double2(akp.x * R.c - t1.x, akp.y * R.c - t1.y);
        a[k + N * R.q] = make_double2(t2.x + akq.x * R.c, t2.y + akq.y * R.c);
        const double2 vkp = v[k + N * R.p], vkq = v[k + N * R.q];
        t1 = cmul(vkq, R.sm);
        t2 = cmul(vkp, R.sp);
        v[k + N * R.p] = make_double2(vkp.x * R.c - t1.x, vkp.y * R.c - t1.y);
        v[k + N * R.q] = make_double2(t2.x + vkq.x * R.c, t2.y + vkq.y * R.c);
      }
      __syncthreads();
      // A <- J^H A (rows p, q of every pair)
      for (int it = li; it < N * NP; it += L) {
        const int k = it % N, pr = it / N;
        const Rot R = rot[pr];
        if (R.q >= N) continue;
        const double2 apk = a[R.p + N * k], aqk = a[R.q + N * k];
        const double2 t1 = cmul(R.sp, aqk), t2 = cmul(R.sm, apk);
        a[R.p + N * k] = make_double2(R.c * apk.x - t1.x, R.c * apk.y - t1.y);
        a[R.q + N * k] = make_double2(t2.x + R.c * aqk.x, t2.y + R.c * aqk.y);
      }
      __syncthreads();
      if (li < NP) {
        const Rot R = rot[li];
        if (R.q < N) {
          a[R.p + N * R.q] = make_double2(0., 0.);
          a[R.q + N * R.p] = make_double2(0., 0.);
          a[R.p + N * R.p].y = 0.;
          a[R.q + N * R.q].y = 0.;
        }
      }
      __syncthreads();
    }
  }
  // ascending order by rank (stable)
  if (li < N) red[li] = a[li + N * li].x;
  __syncthreads();
  if (on && li < N) {
    const int i = li;
    const double di = red[i];
    int rank = 0;
    for (int j = 0; j < N; ++j) {
      const double dj = red[j];
      rank += (dj < di) || (dj == di && j < i);
    }
    double lam = di;
    if (PHONON) lam = lam > 0. ? sqrt(lam) : -sqrt(-lam);
    W[size_t(mat) * N + rank] = lam;
    for (int k = 0; k < N; ++k) {
      double2 x = v[k + N * i];
      if (PHONON) {
        const double s = sqrt(masses[k]);
        x.x /= s;
        x.y /= s;
      }
      Am[k + N * rank] = x;
    }
  }
  if (on && li == 0 && info) info[mat] = mySweeps >= 0 ? mySweeps : -1;
}

template <int N, int L>
static size_t jacobi_n_lds_bytes() {
  constexpr int G = 64 / L, NE = N + (N & 1);
  return size_t(G) * (size_t(2) * N * N * sizeof(double2) + (NE / 2) * sizeof(Rot) + N * sizeof(double));
}

template <int N, int L>
static void launch_jacobi_n(int M, double2* A, double* W, int* info, const double* masses, hipStream_t stream) {
  constexpr int G = 64 / L;
  const size_t lds = jacobi_n_lds_bytes<N, L>();
  const dim3 grid(ceil_div(M, G));
  if (masses)
    hipLaunchKernelGGL((k_jacobi_n<true, N, L>), grid, dim3(64), lds, stream, M, A, W, info, masses, 40);
  else
    hipLaunchKernelGGL((k_jacobi_n<false, N, L>), grid, dim3(64), lds, stream, M, A, W, info, nullptr, 40);
}

// Launch the batched eigensolver on device buffers: the fixed-order kernel
// for the orders of the path's matrices (6, 8, 12, 16 and 18: 2, 4 and 6
// atoms, 8 and 16 Wannier functions), packed 8 or 4 matrices per wave up to
// N = 12 and one per wave above (where 4 would hold 35 KB of LDS per
// workgroup and leave one wave per SIMD), the runtime-order kernel otherwise.
void launch_jacobi(int N, int M, double2* A, double* W, int* info, const double* masses,
                   hipStream_t stream) {
  PHX_REQUIRE(N >= 1 && N <= 64, "eigensolver supports 1 <= N <= 64");
  if (M <= 0) return;
  switch (N) {
    case 6: launch_jacobi_n<6, 8>(M, A, W, info, masses, stream); break;
    case 8: launch_jacobi_n<8, 8>(M, A, W, info, masses, stream); break;
    case 12: launch_jacobi_n<12, 16>(M, A, W, info, masses, stream); break;
    case 16: launch_jacobi_n<16, 64>(M, A, W, info, masses, stream); break;
    case 18: launch_jacobi_n<18, 64>(M, A, W, info, masses, stream); break;
    default: {
      const size_t lds = jacobi_lds_bytes(N);
      if (masses)
        hipLaunchKernelGGL(k_jacobi<true>, dim3(M), dim3(64), lds, stream, N, M, A, W, info, masses, 40);
      else
        hipLaunchKernelGGL(k_jacobi<false>, dim3(M), dim3(64), lds, stream, N, M, A, W, info, nullptr, 40);
    }
  }
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
