// Electron-phonon coupling kernels: InteractionElPhWan::cacheElPh and
// calcCouplingSquared (interaction_elph.cpp:229-448, 492-701) plus the
// Gaussian-delta linewidth reduction of ElScatteringMatrix::builder
// (el_scattering.cpp:216-386), re-designed for gfx950.
//
// Data flow for one batch of outer points k1 (index kk inside the batch):
//
//  k_cache   g1(Rp,nu,w1,w2) = sum_Re e^{ik1.Re}/de gW(Re,Rp,nu,w1,w2)   [k1-batched: every
//            gW element is read once per batch and used for every k1]
//            C(Rp,nu,b1,w2) = sum_w1 g1 U1(w1,b1) * s(k1,Rp)               [fused epilogue]
//            with s = e^{-ik1.Rp}/dp (mesh mode) or 1/dp (explicit-q mode), written
//            as real "panels" B'(k, c), one panel per (k1, b1): rows 2Rp, 2Rp+1 hold
//            C and i*C interleaved, so that a complex product becomes a real GEMM.
//  k_coupling  rows = outer-product partners (k2 or explicit q):
//            g3(row, nu, w2) = sum_Rp A(row, Rp) C(Rp, nu, b1, w2)            [real GEMM on MFMA,
//            A(row, .) = e^{ik2.Rp} interleaved, so e^{iq.Rp} = e^{ik2.Rp} e^{-ik1.Rp}]
//            epilogue in LDS: g4 = e3^T g3 (phonon rotation, e3 not conjugated),
//            g = U2^H g4, |g|^2, then either store |g|^2 (calcCouplingSquared) or
//            symmetrize over degenerate b2 / nu groups and contract with the
//            Gaussian-delta weights into per-workgroup partial linewidths.
//  k_collect  sums the partials in a fixed order (deterministic, no float atomics),
//            including the b1 degenerate-group average of symmetrizeCoupling.
#include <algorithm>
#include <thread>
#include <vector>

#include "phx_common.hpp"

namespace phx {

// ------------------------------------------------------------------ synthetic gW
__host__ __device__ inline uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}
__host__ __device__ inline double hash_u(uint64_t seed, uint64_t ctr) {
  uint64_t h = mix64(seed + ctr * 0x9E3779B97F4A7C15ull);
  return double(h >> 11) * (1.0 / 9007199254740992.0);
}
__host__ __device__ inline double hash_n(uint64_t seed, uint64_t idx) {
  uint64_t c = idx * 4;
  double s = hash_u(seed, c) + hash_u(seed, c + 1);
  s = s + hash_u(seed, c + 2);
  s = s + hash_u(seed, c + 3);
  return (s - 2.0) * 1.7320508075688772;
}

__global__ void k_synth_gw(double2* __restrict__ gW, size_t n, size_t blk,
                           const double* __restrict__ decay, uint64_t seed) {
  for (size_t f = size_t(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += size_t(gridDim.x) * blockDim.x) {
    double d = decay[f / blk];
    gW[f] = make_double2(d * hash_n(seed, 2 * f), d * hash_n(seed, 2 * f + 1));
  }
}

void synth_gw_host(double* gW, size_t n, size_t blk, const double* decay, uint64_t seed) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([=] {
      for (size_t f = t; f < n; f += nt) {
        double d = decay[f / blk];
        gW[2 * f] = d * hash_n(seed, 2 * f);
        gW[2 * f + 1] = d * hash_n(seed, 2 * f + 1);
      }
    });
  for (auto& x : th) x.join();
}

void launch_synth_gw(double2* gW, size_t n, size_t blk, const double* decay, uint64_t seed,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_synth_gw, dim3(8192), dim3(256), 0, s, gW, n, blk, decay, seed);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ small kernels
template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

template <typename T>
void launch_fill(T* p, size_t n, T v, hipStream_t s) {
  if (!n) return;
  int blocks = int(std::min<size_t>(8192, (n + 255) / 256));
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, n, v);
  PHX_HIP(hipGetLastError());
}
template void launch_fill<double>(double*, size_t, double, hipStream_t);
template void launch_fill<float>(float*, size_t, float, hipStream_t);
template void launch_fill<int>(int*, size_t, int, hipStream_t);

// Phase table P(i, r) = e^{sign i k_i.R_r} * scale_r  (double2, rows x nR).
__global__ void k_phase_table(int rows, int nR, const double* __restrict__ kc,
                              const double* __restrict__ R, const double* __restrict__ scale,
                              double sign, double2* __restrict__ P) {
  size_t n = size_t(rows) * nR;
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += size_t(gridDim.x) * blockDim.x) {
    int i = int(t / nR), r = int(t % nR);
    double arg = kc[3 * i] * R[3 * r] + kc[3 * i + 1] * R[3 * r + 1] + kc[3 * i + 2] * R[3 * r + 2];
    double s, c;
    sincos(arg, &s, &c);
    double sc = scale ? scale[r] : 1.0;
    P[t] = make_double2(c * sc, sign * s * sc);
  }
}

void launch_phase_table(int rows, int nR, const double* kc, const double* R, const double* scale,
                        double sign, double2* P, hipStream_t s) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(k_phase_table, dim3(4096), dim3(256), 0, s, rows, nR, kc, R, scale, sign, P);
  PHX_HIP(hipGetLastError());
}

// GEMM A operand: rows of e^{ik.Rp} (no degeneracy factor), interleaved (re, im)
// over Kp >= 2 nRp columns (zero padded), for rowsPad rows (rows >= nrows zero).
template <typename T>
__global__ void k_phase_rows(int nrows, int rowsPad, int nRp, int Kp, const double* __restrict__ kc,
                             const double* __restrict__ R, T* __restrict__ A) {
  size_t n = size_t(rowsPad) * (Kp / 2);
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += size_t(gridDim.x) * blockDim.x) {
    int i = int(t / (Kp / 2)), r = int(t % (Kp / 2));
    double c = 0., s = 0.;
    if (i < nrows && r < nRp) {
      double arg = kc[3 * i] * R[3 * r] + kc[3 * i + 1] * R[3 * r + 1] + kc[3 * i + 2] * R[3 * r + 2];
      sincos(arg, &s, &c);
    }
    A[size_t(i) * Kp + 2 * r] = T(c);
    A[size_t(i) * Kp + 2 * r + 1] = T(s);
  }
}

template <typename T>
void launch_phase_rows(int nrows, int rowsPad, int nRp, int Kp, const double* kc, const double* R,
                       T* A, hipStream_t s) {
  hipLaunchKernelGGL(k_phase_rows<T>, dim3(4096), dim3(256), 0, s, nrows, rowsPad, nRp, Kp, kc, R, A);
  PHX_HIP(hipGetLastError());
}
template void launch_phase_rows<double>(int, int, int, int, const double*, const double*, double*,
                                        hipStream_t);
template void launch_phase_rows<float>(int, int, int, int, const double*, const double*, float*,
                                       hipStream_t);

// Gather windowed eigenvector columns: pool(off_a + w + nw*b) = U(ik_a)(w, lo_a + b).
__global__ void k_gather_eig(int na, int nw, const int* __restrict__ ik, const int* __restrict__ lo,
                             const int* __restrict__ nb, const int* __restrict__ off,
                             const double2* __restrict__ U, double2* __restrict__ pool) {
  int a = blockIdx.x;
  if (a >= na) return;
  int n = nb[a] * nw;
  const double2* src = U + size_t(ik[a]) * nw * nw + size_t(lo[a]) * nw;
  for (int t = threadIdx.x; t < n; t += blockDim.x) pool[off[a] + t] = src[t];
}

void launch_gather_eig(int na, int nw, const int* ik, const int* lo, const int* nb, const int* off,
                       const double2* U, double2* pool, hipStream_t s) {
  if (na <= 0) return;
  hipLaunchKernelGGL(k_gather_eig, dim3(na), dim3(256), 0, s, na, nw, ik, lo, nb, off, U, pool);
  PHX_HIP(hipGetLastError());
}

// Scatter phonon results of a compact q list into full-mesh indexed tables.
__global__ void k_scatter_ph(int nq, int nm, const int* __restrict__ iq, const double* __restrict__ w,
                             const double2* __restrict__ ev, double* __restrict__ wAll,
                             double2* __restrict__ evAll) {
  int j = blockIdx.x;
  if (j >= nq) return;
  int q = iq[j];
  for (int t = threadIdx.x; t < nm * nm; t += blockDim.x)
    evAll[size_t(q) * nm * nm + t] = ev[size_t(j) * nm * nm + t];
  for (int t = threadIdx.x; t < nm; t += blockDim.x) wAll[size_t(q) * nm + t] = w[size_t(j) * nm + t];
}

void launch_scatter_ph(int nq, int nm, const int* iq, const double* w, const double2* ev,
                       double* wAll, double2* evAll, hipStream_t s) {
  if (nq <= 0) return;
  hipLaunchKernelGGL(k_scatter_ph, dim3(nq), dim3(256), 0, s, nq, nm, iq, w, ev, wAll, evAll);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ cacheElPh
constexpr int kCacheRC = 32;  // R_e chunk staged in LDS

struct CacheParams {
  int nRe, nRp, nm, nw, nkk;
  const double2* gW;      // (nRe, nRp, nm, nw, nw)
  const double2* phE;     // (nkk, nRe): e^{ik1.Re}/de
  const double2* phP;     // (nkk, nRp): e^{-ik1.Rp}/dp or 1/dp
  const double2* U1pool;  // per kk: (nw, nb1) column-major
  const int* U1off;       // complex offset per kk
  const int* nb1;         // per kk
  const int* panelBase;   // first panel of kk
  void* B;                // panels
  size_t panelStride;     // elements
  int ldb;                // elements per panel row
};

// One workgroup per (Rp, nu): the (w1, w2) block of gW is streamed once over
// all R_e and contracted with the phases of KK outer points at a time (EPT
// entries of the block per thread).
template <typename OutT, int EPT, int KK>
__global__ void __launch_bounds__(256) k_cache(CacheParams p) {
  __shared__ double2 ph[KK][kCacheRC];
  __shared__ double2 buf[256 * EPT];
  const int rp = blockIdx.x / p.nm, nu = blockIdx.x % p.nm;
  const int E = p.nw * p.nw;
  const int t = threadIdx.x;
  const size_t strideRe = size_t(p.nRp) * p.nm * E;
  const double2* g = p.gW + (size_t(rp) * p.nm + nu) * E;
  OutT* B = static_cast<OutT*>(p.B);
  for (int kk0 = 0; kk0 < p.nkk; kk0 += KK) {
    const int nkk = min(KK, p.nkk - kk0);
    double2 acc[EPT][KK];
#pragma unroll
    for (int u = 0; u < EPT; ++u)
#pragma unroll
      for (int k = 0; k < KK; ++k) acc[u][k] = make_double2(0., 0.);
    for (int r0 = 0; r0 < p.nRe; r0 += kCacheRC) {
      __syncthreads();
      for (int i = t; i < KK * kCacheRC; i += 256) {
        int k = i / kCacheRC, r = i % kCacheRC;
        ph[k][r] = (k < nkk && r0 + r < p.nRe) ? p.phE[size_t(kk0 + k) * p.nRe + r0 + r]
                                               : make_double2(0., 0.);
      }
      __syncthreads();
      const int rmax = min(kCacheRC, p.nRe - r0);
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = t + 256 * u;
        if (e < E) {
          const double2* gp = g + size_t(r0) * strideRe + e;
          for (int r = 0; r < rmax; ++r) {
            double2 x = gp[size_t(r) * strideRe];
#pragma unroll
            for (int k = 0; k < KK; ++k) {
              double2 q = ph[k][r];
              acc[u][k].x += q.x * x.x - q.y * x.y;
              acc[u][k].y += q.x * x.y + q.y * x.x;
            }
          }
        }
      }
    }
    // epilogue: rotate by U1 and scatter into the panels
    for (int k = 0; k < nkk; ++k) {
      const int kk = kk0 + k;
      __syncthreads();
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = t + 256 * u;
        if (e < E) buf[e] = acc[u][k];
      }
      __syncthreads();
      const int nb1 = p.nb1[kk];
      const double2* U1 = p.U1pool + p.U1off[kk];
      const double2 s = p.phP[size_t(kk) * p.nRp + rp];
      for (int o = t; o < nb1 * p.nw; o += 256) {
        const int b1 = o / p.nw, w2 = o % p.nw;
        double2 v = make_double2(0., 0.);
        for (int w1 = 0; w1 < p.nw; ++w1) {
          double2 a = buf[w1 * p.nw + w2], u = U1[w1 + p.nw * b1];
          v.x += a.x * u.x - a.y * u.y;
          v.y += a.x * u.y + a.y * u.x;
        }
        double2 c = make_double2(v.x * s.x - v.y * s.y, v.x * s.y + v.y * s.x);
        OutT* panel = B + size_t(p.panelBase[kk] + b1) * p.panelStride;
        const size_t col = 2 * (size_t(nu) * p.nw + w2);
        panel[size_t(2 * rp) * p.ldb + col] = OutT(c.x);
        panel[size_t(2 * rp) * p.ldb + col + 1] = OutT(c.y);
        panel[size_t(2 * rp + 1) * p.ldb + col] = OutT(-c.y);
        panel[size_t(2 * rp + 1) * p.ldb + col + 1] = OutT(c.x);
      }
    }
  }
}

template <typename OutT>
void launch_cache(const CacheParams& p, hipStream_t s) {
  PHX_REQUIRE(p.nw <= 32, "cacheElPh kernel supports nw <= 32");
  if (p.nkk <= 0) return;
  const int E = p.nw * p.nw;
  if (E <= 256)
    hipLaunchKernelGGL((k_cache<OutT, 1, 16>), dim3(p.nRp * p.nm), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((k_cache<OutT, 4, 4>), dim3(p.nRp * p.nm), dim3(256), 0, s, p);
  PHX_HIP(hipGetLastError());
}
template void launch_cache<double>(const CacheParams&, hipStream_t);
template void launch_cache<float>(const CacheParams&, hipStream_t);

// ------------------------------------------------------------------ coupling
// (coupling_kernel.hpp, instantiated in coupling_f64.hip and coupling_f32.hip)
constexpr int kGMax = 8;
constexpr int kCMax = 4;

// ------------------------------------------------------------------ collect
struct CollectParams {
  int nkk, nMtiles, nCalc, numStates;
  const int* kkNb1; const int* kkPanelBase; const int* kkStateOff;
  const int* panelGStart; const int* panelGSize;
  const double* partial;
  double* linewidth;  // (nCalc, numStates), accumulated into
};

__global__ void k_collect(CollectParams p) {
  const int kk = blockIdx.x;
  if (kk >= p.nkk) return;
  const int nb1 = p.kkNb1[kk];
  for (int i = threadIdx.x; i < nb1 * p.nCalc; i += blockDim.x) {
    const int b1 = i / p.nCalc, c = i % p.nCalc;
    const int pan = p.kkPanelBase[kk] + b1;
    const int gs = p.panelGStart[pan], gz = p.panelGSize[pan];
    double sum = 0.;
    for (int j = 0; j < gz; ++j) {
      const int src = p.kkPanelBase[kk] + gs + j;
      for (int mt = 0; mt < p.nMtiles; ++mt)
        sum += p.partial[(size_t(src) * p.nMtiles + mt) * kGMax * kCMax + (b1 - gs) * kCMax + c];
    }
    p.linewidth[size_t(c) * p.numStates + p.kkStateOff[kk] + b1] += sum;
  }
}

void launch_collect(const CollectParams& p, hipStream_t s) {
  if (p.nkk <= 0) return;
  hipLaunchKernelGGL(k_collect, dim3(p.nkk), dim3(64), 0, s, p);
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
