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

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

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
constexpr int kGMax = 8;  // largest degenerate b1 group handled in one panel
constexpr int kCMax = 4;  // largest number of (T, mu) calculations
constexpr int kNmMax = 36;
constexpr int kKC = 16;   // K chunk staged in LDS

struct CoupParams {
  const void* A; int lda;                 // rows (padded to the row tile) x lda
  const void* B; size_t panelStride; int ldb;
  int M, Kp, nm, nw, nPanels, nMtiles;
  const int* panelKK;
  const int* panelB1;
  // rows
  const double2* U2pool; const int* rowU2off; const int* rowNb2;
  // phonons
  const double2* E3pool; const double* Wpool;
  int qmode;                              // 0: e3 index = rowE3[row]; 1: mesh difference
  const int* rowE3; const int* rowMesh; const int* kkMesh; int m0, m1, m2;
  // store mode (calcCouplingSquared): out(q)(b1 + nb1 (b2 + nb2 nu))
  double* out; const long long* rowOutOff; const int* kkNb1;
  // reduce mode
  const double* e2pool; const int* rowStateOff;
  const double* e1pool; const int* kkStateOff;
  const int* panelGStart; const int* panelGSize;
  const double* fermi; int numStates;
  const double* bose;                     // (q, nu, calc)
  int nCalc; double invSigma, deltaPref, norm, cutoff;
  double* partial;                        // (L = panel*nMtiles + mtile, kGMax, kCMax)
  int nb2max;
};

__device__ inline double gaussian_delta(double energy, double invW, double pref) {
  // GaussianDeltaFunction::getSmearing, delta_function.cpp:40-45
  double x = energy * invW;
  if (x > 6.) return 0.;
  return pref * exp(-x * x);
}

// FP64 / FP32 main loop on v_mfma_*_16x16x4: A lane l holds A[l%16][l/16], B lane l holds
// B[l/16][l%16].  Accumulator i of lane l is D[row(l/16, i)][l%16] with
// row = l/16 + 4i for f64 and 4(l/16) + i for f32 (measured on gfx950).
template <typename T> struct Mfma4;
template <> struct Mfma4<double> {
  typedef d4_t acc_t;
  __device__ static acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int i) { return lk + 4 * i; }
};
template <> struct Mfma4<float> {
  typedef f4_t acc_t;
  __device__ static acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int i) { return 4 * lk + i; }
};

template <typename T, int MT, int NTW, bool REDUCE>
__global__ void __launch_bounds__(256) k_coupling(CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int ROWS = 16 * MT;
  const int nTilesTot = p.nPanels * p.nMtiles;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int L = (bid % 8) * perXcd + bid / 8;  // XCD-contiguous logical index
  if (L >= nTilesTot) return;
  const int panel = L / p.nMtiles, mtile = L % p.nMtiles;
  const int row0 = mtile * ROWS;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int PWp = p.ldb;
  const int ntiles = PWp / 16;

  typedef typename Mfma4<T>::acc_t acc_t;
  acc_t acc[MT][NTW];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int j = 0; j < NTW; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[m][j][i] = 0;

  T* As = reinterpret_cast<T*>(smem);  // ROWS x (kKC + 1)
  const T* A = static_cast<const T*>(p.A) + size_t(row0) * p.lda;
  const T* B = static_cast<const T*>(p.B) + size_t(panel) * p.panelStride;
  const int lr = lane & 15, lk = lane >> 4;

  for (int kc = 0; kc < p.Kp; kc += kKC) {
    __syncthreads();
    for (int i = t; i < ROWS * kKC; i += 256) {
      int r = i / kKC, k = i % kKC;
      As[r * (kKC + 1) + k] = A[size_t(r) * p.lda + kc + k];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kKC / 4; ++ks) {
      T a[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) a[m] = As[(16 * m + lr) * (kKC + 1) + ks * 4 + lk];
      const T* Bk = B + size_t(kc + ks * 4 + lk) * PWp + lr;
#pragma unroll
      for (int j = 0; j < NTW; ++j) {
        const int tile = wave + 4 * j;
        if (tile < ntiles) {
          T b = Bk[tile * 16];
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[m][j] = Mfma4<T>::mac(a[m], b, acc[m][j]);
        }
      }
    }
  }

  // ---------------- epilogue: g3 tile into LDS (row-major ROWS x PWp, type T)
  __syncthreads();
  T* g3 = reinterpret_cast<T*>(smem);
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
      const int tile = wave + 4 * j;
      if (tile < ntiles) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          g3[(16 * m + Mfma4<T>::row(lk, i)) * PWp + tile * 16 + lr] = acc[m][j][i];
      }
    }
  __syncthreads();

  const int nm = p.nm, nw = p.nw;
  constexpr int TPR = 256 / ROWS;  // threads per row
  const int r = t / TPR, sub = t % TPR;
  const int row = row0 + r;
  const bool valid = row < p.M;
  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];
  int iq = 0;
  if (valid) iq = p.qmode ? mesh_diff_index(p.rowMesh[row], p.kkMesh[kk], p.m0, p.m1, p.m2) : p.rowE3[row];
  const double2* e3 = p.E3pool + size_t(iq) * nm * nm;
  T* g = g3 + size_t(r) * PWp;

  // g4(nu2, w2) = sum_nu e3(nu, nu2) g3(nu, w2)  (interaction_elph.cpp:381-390), in place per w2
  if (valid) {
    for (int w2 = sub; w2 < nw; w2 += TPR) {
      double2 x[kNmMax];
#pragma unroll
      for (int nu = 0; nu < kNmMax; ++nu)
        if (nu < nm) x[nu] = make_double2(double(g[2 * (nu * nw + w2)]), double(g[2 * (nu * nw + w2) + 1]));
      for (int nu2 = 0; nu2 < nm; ++nu2) {
        double2 s = make_double2(0., 0.);
#pragma unroll
        for (int nu = 0; nu < kNmMax; ++nu)
          if (nu < nm) {
            double2 e = e3[nu + nm * nu2];
            s.x += x[nu].x * e.x - x[nu].y * e.y;
            s.y += x[nu].x * e.y + x[nu].y * e.x;
          }
        g[2 * (nu2 * nw + w2)] = T(s.x);
        g[2 * (nu2 * nw + w2) + 1] = T(s.y);
      }
    }
  }
  __syncthreads();

  // g(nu, b1, b2) = sum_w2 conj(U2(w2, b2)) g4(nu, w2)  (interaction_elph.cpp:393-402)
  const int nb2 = valid ? p.rowNb2[row] : 0;
  const double2* U2 = p.U2pool + (valid ? p.rowU2off[row] : 0);
  double* gsq = reinterpret_cast<double*>(g3 + size_t(ROWS) * PWp);  // ROWS x (nm * nb2max)
  const int gstride = nm * p.nb2max;
  double* gs = gsq + size_t(r) * gstride;
  for (int o = sub; o < nm * nb2; o += TPR) {
    const int nu = o / nb2, b2 = o % nb2;
    double2 s = make_double2(0., 0.);
    for (int w2 = 0; w2 < nw; ++w2) {
      double2 u = U2[w2 + nw * b2];
      double2 x = make_double2(double(g[2 * (nu * nw + w2)]), double(g[2 * (nu * nw + w2) + 1]));
      s.x += u.x * x.x + u.y * x.y;
      s.y += u.x * x.y - u.y * x.x;
    }
    const double v = s.x * s.x + s.y * s.y;
    if (REDUCE) {
      gs[nu * nb2 + b2] = v;
    } else {
      const int nb1 = p.kkNb1[kk];
      p.out[p.rowOutOff[row] + b1p + size_t(nb1) * (b2 + size_t(nb2) * nu)] = v;
    }
  }
  if (!REDUCE) return;
  __syncthreads();

  const double* w3 = p.Wpool + size_t(iq) * nm;
  const double* e2 = p.e2pool + (valid ? p.rowStateOff[row] : 0);
  // symmetrizeCoupling (scattering.cpp:1209-1318) along b2, then nu
  for (int nu = sub; nu < nm && valid; nu += TPR) {
    for (int b = 0; b < nb2; ++b) {
      int deg = 0;
      for (int i = b; i < nb2; ++i) {
        if (fabs(e2[b] - e2[i]) < 1.0e-6) deg++;
        else break;
      }
      if (deg > 1) {
        double s = 0.;
        for (int i = 0; i < deg; ++i) s += gs[nu * nb2 + b + i] / double(float(deg));
        for (int i = 0; i < deg; ++i) gs[nu * nb2 + b + i] = s;
      }
      b += deg - 1;
    }
  }
  __syncthreads();
  for (int b2 = sub; b2 < nb2; b2 += TPR) {
    for (int nu = 0; nu < nm; ++nu) {
      int deg = 0;
      for (int i = nu; i < nm; ++i) {
        if (fabs(w3[nu] - w3[i]) < 1.0e-6) deg++;
        else break;
      }
      if (deg > 1) {
        double s = 0.;
        for (int i = 0; i < deg; ++i) s += gs[(nu + i) * nb2 + b2] / double(float(deg));
        for (int i = 0; i < deg; ++i) gs[(nu + i) * nb2 + b2] = s;
      }
      nu += deg - 1;
    }
  }
  __syncthreads();

  // Gaussian-delta weights (el_scattering.cpp:277-380, switchCase 2)
  const int gstart = p.panelGStart[panel], gsize = p.panelGSize[panel];
  const double* e1 = p.e1pool + p.kkStateOff[kk];
  const int nCalc = p.nCalc;
  double part[kGMax][kCMax];
#pragma unroll
  for (int gi = 0; gi < kGMax; ++gi)
#pragma unroll
    for (int c = 0; c < kCMax; ++c) part[gi][c] = 0.;
  const double inv_g = 1. / double(float(gsize));
  for (int o = sub; o < nm * nb2; o += TPR) {
    const int nu = o / nb2, b2 = o % nb2;
    const double en3 = w3[nu];
    if (en3 < p.cutoff) continue;
    const double cpl = gs[nu * nb2 + b2] * inv_g;
    const double en2 = e2[b2];
    const int is2 = p.rowStateOff[row] + b2;
#pragma unroll
    for (int gi = 0; gi < kGMax; ++gi) {
      if (gi >= gsize) break;
      const double en1 = e1[gstart + gi];
      const double d1 = gaussian_delta(en1 - en2 + en3, p.invSigma, p.deltaPref);
      const double d2 = gaussian_delta(en1 - en2 - en3, p.invSigma, p.deltaPref);
      if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
      for (int c = 0; c < kCMax; ++c) {
        if (c >= nCalc) break;
        const double f2 = p.fermi[size_t(c) * p.numStates + is2];
        const double n3 = p.bose[(size_t(iq) * nm + nu) * nCalc + c];
        part[gi][c] += cpl * ((f2 + n3) * d1 + (1. - f2 + n3) * d2) * p.norm / en3 * 3.14159265358979323846;
      }
    }
  }
  // block reduction of part -> partial[L]
  double* red = reinterpret_cast<double*>(smem);  // reuse (after the barrier)
  __syncthreads();
#pragma unroll
  for (int gi = 0; gi < kGMax; ++gi)
#pragma unroll
    for (int c = 0; c < kCMax; ++c) {
      double v = part[gi][c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[(gi * kCMax + c) * 4 + wave] = v;
    }
  __syncthreads();
  if (t < kGMax * kCMax) {
    double v = red[t * 4] + red[t * 4 + 1] + red[t * 4 + 2] + red[t * 4 + 3];
    p.partial[size_t(L) * kGMax * kCMax + t] = v;
  }
}

size_t coupling_lds_bytes(int elem, int MT, int ldb, int nm, int nb2max, bool reduce) {
  const int ROWS = 16 * MT;
  size_t a = size_t(ROWS) * (kKC + 1) * elem;
  size_t e = size_t(ROWS) * ldb * elem + (reduce ? size_t(ROWS) * nm * nb2max * sizeof(double) : 0);
  size_t r = size_t(kGMax) * kCMax * 4 * sizeof(double);
  return std::max(std::max(a, e), r);
}

template <typename T, int MT, int NTW, bool REDUCE>
static void launch_coupling_t(const CoupParams& p, int grid, size_t lds, hipStream_t s) {
  auto kern = k_coupling<T, MT, NTW, REDUCE>;
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, p);
  PHX_HIP(hipGetLastError());
}

template <typename T, int MT, bool REDUCE>
static void dispatch_ntw(int ntw, const CoupParams& p, int grid, size_t lds, hipStream_t s) {
  switch (ntw) {
#define PHX_NTW(n) case n: launch_coupling_t<T, MT, n, REDUCE>(p, grid, lds, s); break;
    PHX_NTW(1) PHX_NTW(2) PHX_NTW(3) PHX_NTW(4) PHX_NTW(5) PHX_NTW(6) PHX_NTW(7) PHX_NTW(8) PHX_NTW(9)
#undef PHX_NTW
    default: throw Error(-1, "coupling kernel: panel too wide");
  }
}

// rowsPerTile() is 32 for panels up to 384 real columns, else 16.
int coupling_rows_per_tile(int ldb) { return ldb <= 384 ? 32 : 16; }

template <typename T>
void launch_coupling(const CoupParams& p0, bool reduce, hipStream_t s) {
  CoupParams p = p0;
  PHX_REQUIRE(p.nm <= kNmMax, "coupling kernel supports nm <= 36");
  PHX_REQUIRE(p.ldb % 16 == 0 && p.Kp % kKC == 0, "coupling kernel: bad padding");
  const int MT = coupling_rows_per_tile(p.ldb) / 16;
  const int ntw = ceil_div(p.ldb / 16, 4);
  const int tiles = p.nPanels * p.nMtiles;
  if (tiles <= 0) return;
  const int grid = ceil_div(tiles, 8) * 8;
  const size_t lds = coupling_lds_bytes(sizeof(T), MT, p.ldb, p.nm, p.nb2max, reduce);
  PHX_REQUIRE(lds <= 160 * 1024, "coupling kernel: LDS budget exceeded");
  if (MT == 2) {
    if (reduce) dispatch_ntw<T, 2, true>(ntw, p, grid, lds, s);
    else dispatch_ntw<T, 2, false>(ntw, p, grid, lds, s);
  } else {
    if (reduce) dispatch_ntw<T, 1, true>(ntw, p, grid, lds, s);
    else dispatch_ntw<T, 1, false>(ntw, p, grid, lds, s);
  }
}
template void launch_coupling<double>(const CoupParams&, bool, hipStream_t);
template void launch_coupling<float>(const CoupParams&, bool, hipStream_t);

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
