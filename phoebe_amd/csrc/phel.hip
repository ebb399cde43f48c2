// Phonon linewidths from electron-phonon scattering on gfx950
// (PhElScatteringMatrix::builder, phel_scattering.cpp:43-363), the reduction
// half: for one outer electron point k1 and all phonon wavevectors q3 of the
// mesh, the couplings |g(b1, b2, nu)|^2 at k2 = k1 + q3 (calcCouplingSquared in
// store mode, (b1, b2, nu) column-major per q3) are
//   * symmetrized over degenerate b1, then b2, then nu groups
//     (ScatteringMatrix::symmetrizeCoupling, scattering.cpp:1209-1318), and
//   * reduced into the phonon linewidths
//     Gamma(c, q3, nu) += sum_{b1, b2} |g|^2 F_c(b1) delta(e1 - e2 + w) norm / T_c * pi
//     with F = f (1 - f) = 1 / (4 cosh^2((e - mu) / 2T)) (phel_scattering.cpp:284-324).
// Every (q3, nu, c) entry is owned by one thread and gains one term per k1
// launch, so the sums are deterministic (launch order = k1 order).
#include "coupling_kernel.hpp"
#include "phel.hpp"
#include "phx_common.hpp"

namespace phx {


__global__ void __launch_bounds__(64) k_phel_sym(PhelParams p) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.nq) return;
  const int nb1 = p.nb1, nb2 = p.nb2, nm = p.nm;
  double* c = p.g2 + size_t(q) * p.g2Stride;
  auto C = [&](int a, int b, int v) -> double& { return c[a + size_t(nb1) * (b + size_t(nb2) * v)]; };
  const double* e2 = p.E2 + p.e2off[q];
  const double* w3 = p.W3 + size_t(q) * nm;
  for (int a = 0; a < nb1;) {
    int deg = 0;
    for (int i = a; i < nb1; ++i) {
      if (fabs(p.e1[a] - p.e1[i]) < 1.0e-6) ++deg;
      else break;
    }
    if (deg > 1)
      for (int b = 0; b < nb2; ++b)
        for (int v = 0; v < nm; ++v) {
          double s = 0.;
          for (int i = 0; i < deg; ++i) s += C(a + i, b, v) / double(float(deg));
          for (int i = 0; i < deg; ++i) C(a + i, b, v) = s;
        }
    a += deg;
  }
  for (int b = 0; b < nb2;) {
    int deg = 0;
    for (int i = b; i < nb2; ++i) {
      if (fabs(e2[b] - e2[i]) < 1.0e-6) ++deg;
      else break;
    }
    if (deg > 1)
      for (int a = 0; a < nb1; ++a)
        for (int v = 0; v < nm; ++v) {
          double s = 0.;
          for (int i = 0; i < deg; ++i) s += C(a, b + i, v) / double(float(deg));
          for (int i = 0; i < deg; ++i) C(a, b + i, v) = s;
        }
    b += deg;
  }
  for (int v = 0; v < nm;) {
    int deg = 0;
    for (int i = v; i < nm; ++i) {
      if (fabs(w3[v] - w3[i]) < 1.0e-6) ++deg;
      else break;
    }
    if (deg > 1)
      for (int a = 0; a < nb1; ++a)
        for (int b = 0; b < nb2; ++b) {
          double s = 0.;
          for (int i = 0; i < deg; ++i) s += C(a, b, v + i) / double(float(deg));
          for (int i = 0; i < deg; ++i) C(a, b, v + i) = s;
        }
    v += deg;
  }
}

__device__ inline double phel_delta(const PhelParams& p, double energy, int q, int a, int b) {
  if (!p.adaptive) return gaussian_delta(energy, p.invSigma, p.deltaPref);
  const double* x = p.v1 + 3 * a;
  const double* y = p.V2 + (size_t(q) * p.nb2 + b) * 3;
  const double v0 = x[0] - y[0], v1 = x[1] - y[1], v2 = x[2] - y[2];
  if (v0 == 0. && v1 == 0. && v2 == 0.) return energy == 0. ? 1. : 0.;
  double s = 0.;
  for (int i = 0; i < 3; ++i) {
    const double d = p.qT[3 * i] * v0 + p.qT[3 * i + 1] * v1 + p.qT[3 * i + 2] * v2;
    s += d * d;
  }
  s = sqrt(s / 6.);
  if (s == 0.) return 0.;
  if (s < p.minWidth) s = p.minWidth;
  if (fabs(energy) > 2. * s) return 0.;
  const double t = energy / s;
  if (t > 6.) return 0.;
  return exp(-t * t) / 1.7724538509055160273 / s / 0.9953222650189527;
}

__global__ void __launch_bounds__(256) k_phel_rates(PhelParams p) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.nq * p.nm) return;
  const int q = t / p.nm, v = t % p.nm;
  const double w = p.W3[size_t(q) * p.nm + v];
  if (w < p.cutoff) return;  // phel_scattering.cpp:266-270
  const int nb1 = p.nb1, nb2 = p.nb2;
  const double* c = p.g2 + size_t(q) * p.g2Stride;
  const double* e2 = p.E2 + p.e2off[q];
  double acc[kCMax];
  for (int k = 0; k < kCMax; ++k) acc[k] = 0.;
  for (int a = 0; a < nb1; ++a)
    for (int b = 0; b < nb2; ++b) {
      const double d = phel_delta(p, p.e1[a] - e2[b] + w, q, a, b);
      if (!(d > 0.)) continue;
      const double g = c[a + size_t(nb1) * (b + size_t(nb2) * v)];
      for (int k = 0; k < kCMax; ++k)
        if (k < p.nCalc) acc[k] += g * p.F1[size_t(k) * nb1 + a] * d * p.norm * p.invT[k] * 3.14159265358979323846;
    }
  for (int k = 0; k < kCMax; ++k)
    if (k < p.nCalc) p.lw[(size_t(k) * p.nq + q) * p.nm + v] += acc[k];
}

// flag[0] = 1 when any eigensolve status in info[0..n) is negative
__global__ void __launch_bounds__(256) k_phel_flag_failures(int n, const int* info, int* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && info[i] < 0) flag[0] = 1;
}

void launch_phel_flag_failures(int n, const int* info, int* flag, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_phel_flag_failures, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, info, flag);
  PHX_HIP(hipGetLastError());
}

void launch_phel_reduce(const PhelParams& p, hipStream_t s) {
  if (p.nq <= 0) return;
  if (p.nCalc < 1 || p.nCalc > kCMax) throw Error(-1, "ph-el linewidths: 1..4 (T, mu) calculations");
  hipLaunchKernelGGL(k_phel_sym, dim3(ceil_div(p.nq, 64)), dim3(64), 0, s, p);
  PHX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_phel_rates, dim3(ceil_div((long long)p.nq * p.nm, 256)), dim3(256), 0, s, p);
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
