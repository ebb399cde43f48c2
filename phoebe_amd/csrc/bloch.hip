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
