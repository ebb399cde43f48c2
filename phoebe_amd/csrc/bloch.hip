// Batched Fourier transforms of the harmonic Hamiltonians.
//
//  * electrons: H(k)_{mn} = sum_R e^{+ik.R}/d(R) h0R(R,m,n)
//    (ElectronH0Wannier::kokkosBatchedBuildBlochHamiltonian,
//     electron_h0_wannier.cpp:777-821);
//  * phonons:   D(q)_{mn} = sum_R e^{-iq.R} w(R) Phi(R,m,n)
//    (PhononH0::kokkosBatchedBuildBlochHamiltonian, phonon_h0_kokkos.cpp:6-46);
//    hermitization and mass scaling are fused into the eigensolver prologue.
//
// Each workgroup owns TK wavevectors and a slice of the matrix entries; the
// phases of a chunk of R vectors are generated once into LDS (FP64 sincos)
// and every thread streams its entry's coefficient column with coalesced
// loads, accumulating TK complex sums in registers.
#include "phx_common.hpp"

namespace phx {

constexpr int kBlochTK = 16;
constexpr int kBlochRC = 32;

template <bool PHONON>
__global__ void __launch_bounds__(256) k_bloch(int nk, int nE, int nR,
                                               const double* __restrict__ kc,  // nk x 3
                                               const double* __restrict__ R,   // nR x 3
                                               const double* __restrict__ rw,  // nR (d or w)
                                               const void* __restrict__ coef,  // (nR, nE)
                                               double2* __restrict__ out,      // (nk, nE) col-major per k
                                               int nw) {
  __shared__ double2 ph[kBlochTK][kBlochRC];
  const int k0 = blockIdx.x * kBlochTK;
  const int e = blockIdx.y * 256 + threadIdx.x;
  double2 acc[kBlochTK];
#pragma unroll
  for (int t = 0; t < kBlochTK; ++t) acc[t] = make_double2(0., 0.);
  for (int r0 = 0; r0 < nR; r0 += kBlochRC) {
    __syncthreads();
    for (int i = threadIdx.x; i < kBlochTK * kBlochRC; i += 256) {
      int t = i / kBlochRC, r = i % kBlochRC;
      int k = k0 + t, iR = r0 + r;
      double2 p = make_double2(0., 0.);
      if (k < nk && iR < nR) {
        double arg = kc[3 * k] * R[3 * iR] + kc[3 * k + 1] * R[3 * iR + 1] +
                     kc[3 * k + 2] * R[3 * iR + 2];
        double s, c;
        sincos(arg, &s, &c);
        if (PHONON) p = make_double2(c * rw[iR], -s * rw[iR]);
        else p = make_double2(c / rw[iR], s / rw[iR]);
      }
      ph[t][r] = p;
    }
    __syncthreads();
    if (e < nE) {
      const int rmax = min(kBlochRC, nR - r0);
      for (int r = 0; r < rmax; ++r) {
        double2 h;
        if (PHONON) h = make_double2(static_cast<const double*>(coef)[size_t(r0 + r) * nE + e], 0.);
        else h = static_cast<const double2*>(coef)[size_t(r0 + r) * nE + e];
#pragma unroll
        for (int t = 0; t < kBlochTK; ++t) {
          double2 p = ph[t][r];
          acc[t].x += p.x * h.x - p.y * h.y;
          acc[t].y += p.x * h.y + p.y * h.x;
        }
      }
    }
  }
  if (e < nE) {
    // coefficient entry e = m*nw + n (row-major (m,n)); output column-major
    const int m = e / nw, n = e % nw;
#pragma unroll
    for (int t = 0; t < kBlochTK; ++t) {
      int k = k0 + t;
      if (k < nk) out[size_t(k) * nE + m + size_t(nw) * n] = acc[t];
    }
  }
}

// Rigid-ion long-range term added to D(q) before the eigensolve
// (PhononH0::kokkosBatchedBuildBlochHamiltonian, phonon_h0_kokkos.cpp:48-149,
// the Kokkos form of QE's rgd_blk):
//   D(a i, b j) += c(a, i, j) [a = b]
//                + sum_G n(k) e^{i (tau_a - tau_b).k} Z(a, i; k) Z(b, j; k),  k = G + q,
//   Z(a, i; k) = sum_j k_j born(a, j, i),  n(k) = norm e^{-k.eps.k / 4} / (k.eps.k)
//   for 0 < k.eps.k < 4 gMax (other G contribute nothing).
// Per G this is a rank-one update D += n_G v_G v_G^H with
// v_G(a i) = e^{i tau_a.k} Z(a, i; k).  One workgroup per q: 64 G vectors at a
// time put n_G and v_G in LDS, then every thread accumulates its entries.
constexpr int kLrChunk = 64;
constexpr int kLrMaxModes = 36;  // 12 atoms
constexpr int kLrPer = (kLrMaxModes * kLrMaxModes + 255) / 256;

__global__ void __launch_bounds__(256) k_ph_long_range(int nq, const double* __restrict__ qc,
                                                       LongRange lr, double2* __restrict__ D) {
  __shared__ double2 v[kLrChunk][kLrMaxModes];
  __shared__ double wgt[kLrChunk];
  const int q = blockIdx.x;
  const int nm = 3 * lr.nat, nE = nm * nm;
  const double q0 = qc[3 * q], q1 = qc[3 * q + 1], q2 = qc[3 * q + 2];
  double2 acc[kLrPer];
#pragma unroll
  for (int t = 0; t < kLrPer; ++t) acc[t] = make_double2(0., 0.);
  for (int g0 = 0; g0 < lr.nG; g0 += kLrChunk) {
    __syncthreads();
    if (threadIdx.x < kLrChunk) {
      const int g = g0 + int(threadIdx.x);
      double k0 = 0., k1 = 0., k2 = 0., w = 0.;
      if (g < lr.nG) {
        k0 = lr.G[3 * g] + q0;
        k1 = lr.G[3 * g + 1] + q1;
        k2 = lr.G[3 * g + 2] + q2;
        const double kk[3] = {k0, k1, k2};
        double geg = 0.;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) geg += kk[i] * lr.eps[3 * i + j] * kk[j];
        if (geg > 0. && geg < 4. * lr.gMax) w = lr.norm * exp(-geg * 0.25) / geg;
      }
      wgt[threadIdx.x] = w;
      for (int a = 0; a < lr.nat; ++a) {
        const double arg = lr.tau[3 * a] * k0 + lr.tau[3 * a + 1] * k1 + lr.tau[3 * a + 2] * k2;
        double sn, cs;
        sincos(arg, &sn, &cs);
        for (int i = 0; i < 3; ++i) {
          const double z = k0 * lr.born[9 * a + i] + k1 * lr.born[9 * a + 3 + i] + k2 * lr.born[9 * a + 6 + i];
          v[threadIdx.x][3 * a + i] = make_double2(cs * z, sn * z);
        }
      }
    }
    __syncthreads();
    const int cnt = min(kLrChunk, lr.nG - g0);
#pragma unroll
    for (int t = 0; t < kLrPer; ++t) {
      const int e = int(threadIdx.x) + 256 * t;
      if (e < nE) {
        const int m = e % nm, n = e / nm;
        double2 a = acc[t];
        for (int g = 0; g < cnt; ++g) {
          const double w = wgt[g];
          const double2 x = v[g][m], y = v[g][n];
          a.x += w * (x.x * y.x + x.y * y.y);  // w x conj(y)
          a.y += w * (x.y * y.x - x.x * y.y);
        }
        acc[t] = a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kLrPer; ++t) {
    const int e = int(threadIdx.x) + 256 * t;
    if (e < nE) {
      const int m = e % nm, n = e / nm;  // column-major entry (m, n) at e
      double2 d = D[size_t(q) * nE + e];
      d.x += acc[t].x;
      d.y += acc[t].y;
      if (m / 3 == n / 3) d.x += lr.cst[9 * (m / 3) + 3 * (m % 3) + (n % 3)];
      D[size_t(q) * nE + e] = d;
    }
  }
}

void launch_ph_long_range(int nq, const double* qc, const LongRange& lr, double2* D,
                          hipStream_t stream) {
  if (nq <= 0 || lr.nat <= 0) return;
  if (lr.nat * 3 > kLrMaxModes) throw Error(-1, "long-range term: at most 12 atoms");
  hipLaunchKernelGGL(k_ph_long_range, dim3(nq), dim3(256), 0, stream, nq, qc, lr, D);
  PHX_HIP(hipGetLastError());
}

void launch_bloch(bool phonon, int nk, int nw, int nR, const double* kc, const double* R,
                  const double* rw, const void* coef, double2* out, hipStream_t stream) {
  if (nk <= 0) return;
  const int nE = nw * nw;
  dim3 grid(ceil_div(nk, kBlochTK), ceil_div(nE, 256));
  if (phonon)
    hipLaunchKernelGGL(k_bloch<true>, grid, dim3(256), 0, stream, nk, nE, nR, kc, R, rw, coef, out, nw);
  else
    hipLaunchKernelGGL(k_bloch<false>, grid, dim3(256), 0, stream, nk, nE, nR, kc, R, rw, coef, out, nw);
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
