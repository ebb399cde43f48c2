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

#include "elph_params.hpp"
#include "phx_common.hpp"

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

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

template <typename T2>
__global__ void k_synth_gw(T2* __restrict__ gW, size_t n, size_t blk,
                           const double* __restrict__ decay, uint64_t seed) {
  for (size_t f = size_t(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += size_t(gridDim.x) * blockDim.x) {
    double d = decay[f / blk];
    T2 v;
    v.x = d * hash_n(seed, 2 * f);
    v.y = d * hash_n(seed, 2 * f + 1);
    gW[f] = v;
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
  hipLaunchKernelGGL(k_synth_gw<double2>, dim3(8192), dim3(256), 0, s, gW, n, blk, decay, seed);
  PHX_HIP(hipGetLastError());
}

// FP32 copy of the same tensor, each element rounded once from its FP64 value.
void launch_synth_gw_f32(float2* gW, size_t n, size_t blk, const double* decay, uint64_t seed,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_synth_gw<float2>, dim3(8192), dim3(256), 0, s, gW, n, blk, decay, seed);
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

// Phonon eigenvectors of every q in the layout the separable transform's
// epilogue reads, in its precision: out(q)(nu, nu2) at nu * nm + nu2, i.e.
// the lanes that serve the modes nu2 of one row read consecutive elements.
template <typename T2v>
__global__ void k_transpose_e3(int nk, int nm, const double2* __restrict__ ev, T2v* __restrict__ out) {
  const int q = blockIdx.x;
  if (q >= nk) return;
  const size_t base = size_t(q) * nm * nm;
  for (int t = threadIdx.x; t < nm * nm; t += blockDim.x) {
    const int nu = t / nm, nu2 = t % nm;
    const double2 e = ev[base + nu + size_t(nm) * nu2];
    out[base + t] = T2v{decltype(out->x)(e.x), decltype(out->x)(e.y)};
  }
}

void launch_transpose_e3(int nk, int nm, const double2* ev, void* out, bool fp32, hipStream_t s) {
  if (nk <= 0) return;
  if (fp32)
    hipLaunchKernelGGL(k_transpose_e3<float2>, dim3(nk), dim3(256), 0, s, nk, nm, ev,
                       static_cast<float2*>(out));
  else
    hipLaunchKernelGGL(k_transpose_e3<double2>, dim3(nk), dim3(256), 0, s, nk, nm, ev,
                       static_cast<double2*>(out));
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ cacheElPh
// g1(k1; Rp,nu,w1,w2) = sum_Re P(k1,Re) gW(Re,Rp,nu,w1,w2) as a real MFMA GEMM.
// Output rows m < NKK hold Re(g1) of outer point m, rows NKK+m hold Im(g1);
// the K dimension walks R_e in chunks of 8 vectors, ordered so that lane
// group g of k-step ks reads (Re, Im) of gW at R_e = re0 + g + 4*(ks/2) with
// one 2-element load serving k-steps 2q (Re part) and 2q+1 (Im part):
//   A'(m, k) = [ Pr | -Pi ] for m < NKK,  [ Pi | Pr ] for m >= NKK.
// One workgroup owns one (Rp, nu) block of nw^2 columns (256 per pass), so the
// U1 rotation and the e^{-ik1.Rp}/dp factor run as its epilogue.

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

template <typename T> struct MfmaC;
template <> struct MfmaC<double> {
  typedef d4_t acc_t;
  __device__ static acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int i) { return lk + 4 * i; }
};
template <> struct MfmaC<float> {
  typedef f4_t acc_t;
  __device__ static acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int i) { return 4 * lk + i; }
};

// Column passes of the cacheElPh kernels cover whole w1 blocks: (256 / nw) * nw columns.
__host__ __device__ inline int cache_pass_cols(int nw) { return (256 / nw) * nw; }

// Panel epilogue shared by the cacheElPh kernels.  Cb holds g1 of up to 16
// outer points (kbase .. min(kbase+16, kend)-1) for the columns c0 .. c0+255 of block (rp, nu),
// of which the pass owns c0 .. c0 + cache_pass_cols(nw) - 1 (whole w1 blocks):
// real parts in rows 0..15, imaginary parts in rows 16..31.  Applies the U1
// rotation over the w1 range of those columns and the factor s(k1, Rp), and
// writes (c0 == 0) or adds (later column passes) the panel entries.  Each
// output (outer point, b1, w2) is owned by the same thread in every pass (the
// thread-to-output map does not depend on c0).
template <typename T, typename OutT>
__device__ __forceinline__ void cache_panel_epilogue(const CacheParams& p, const T* Cb, int c0,
                                                     int kbase, int rp, int nu, int kend) {
  const int t = threadIdx.x, nw = p.nw;
  OutT* Bout = static_cast<OutT*>(p.B);
  const int nkl = min(16, kend - kbase);
  // one thread per (outer point, b1, w2) of the batch: the band count of the
  // batch's widest point bounds b1, so all 256 threads share the work
  int nbm = 0;
  for (int kl = 0; kl < nkl; ++kl) nbm = max(nbm, p.nb1[kbase + kl]);
  nbm = __builtin_amdgcn_readfirstlane(nbm);
  const int per = nbm * nw;
  // columns of this pass: w1 in [c0/nw, (c0 + pass)/nw), all w2
  const int w1lo = c0 / nw, w1hi = min(nw, (c0 + cache_pass_cols(nw)) / nw);
  for (int idx = t; idx < nkl * per; idx += 256) {
    const int kl = idx / per, rem = idx - kl * per, b1 = rem / nw, w2 = rem - b1 * nw;
    const int kk = kbase + kl;
    if (b1 >= p.nb1[kk]) continue;
    const double2* U1 = p.U1pool + p.U1off[kk];
    const double2 s = p.phP[size_t(kk) * p.nRp + rp];
    double2 v = make_double2(0., 0.);
    for (int w1 = w1lo; w1 < w1hi; ++w1) {
      const int cl = w1 * nw + w2 - c0;
      const double ar = double(Cb[kl * 256 + cl]), ai = double(Cb[(16 + kl) * 256 + cl]);
      const double2 u = U1[w1 + nw * b1];
      v.x += ar * u.x - ai * u.y;
      v.y += ar * u.y + ai * u.x;
    }
    const double2 c = make_double2(v.x * s.x - v.y * s.y, v.x * s.y + v.y * s.x);
    OutT* panel = Bout + size_t(p.panelBase[kk] + b1) * p.panelStride;
    const size_t col = 2 * (size_t(nu) * nw + w2);
    if (c0 == 0) {
      panel[size_t(2 * rp) * p.ldb + col] = OutT(c.x);
      panel[size_t(2 * rp) * p.ldb + col + 1] = OutT(c.y);
      panel[size_t(2 * rp + 1) * p.ldb + col] = OutT(-c.y);
      panel[size_t(2 * rp + 1) * p.ldb + col + 1] = OutT(c.x);
    } else {  // later column passes add their w1 range
      panel[size_t(2 * rp) * p.ldb + col] += OutT(c.x);
      panel[size_t(2 * rp) * p.ldb + col + 1] += OutT(c.y);
      panel[size_t(2 * rp + 1) * p.ldb + col] += OutT(-c.y);
      panel[size_t(2 * rp + 1) * p.ldb + col + 1] += OutT(c.x);
    }
  }
}

template <typename T, typename OutT>
__global__ void __launch_bounds__(256, 2) k_cache(CacheParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NKK = kCacheNKK, MR = 2 * NKK, MTc = MR / 16, LDA = 17;
  typedef typename Vec2<T>::type T2;
  typedef typename MfmaC<T>::acc_t acc_t;
  const int rp = blockIdx.x / p.nm, nu = blockIdx.x % p.nm;
  const int E = p.nw * p.nw;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15, lg = lane >> 4;
  const size_t strideRe = size_t(p.nRp) * p.nm * E;
  const T2* G = static_cast<const T2*>(p.gW) + (size_t(rp) * p.nm + nu) * E;
  T* As = reinterpret_cast<T*>(smem);               // MR x LDA
  T* Cb = reinterpret_cast<T*>(smem);               // epilogue: 2 x 16 x 256
  const int nchunk = (p.nRe + 7) / 8;

  for (int c0 = 0; c0 < E; c0 += cache_pass_cols(p.nw)) {
    acc_t acc[MTc][4];
#pragma unroll
    for (int m = 0; m < MTc; ++m)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[m][j][i] = 0;
    T2 bq[2][4], bn[2][4];
    auto load_b = [&](T2(&b)[2][4], int re0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int col = c0 + (wave * 4 + j) * 16 + lr;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int re = re0 + lg + 4 * h;
          b[h][j] = (re < p.nRe && col < E) ? G[size_t(re) * strideRe + col] : T2{T(0), T(0)};
        }
      }
    };
    load_b(bq, 0);
    for (int ch = 0; ch < nchunk; ++ch) {
      const int re0 = ch * 8;
      __syncthreads();
      for (int i = t; i < MR * 16; i += 256) {
        const int m = i / 16, k = i % 16, ks = k / 4, g = k % 4;
        const int re = re0 + g + 4 * (ks >> 1), s = ks & 1, kk = m % NKK;
        const double2 P = (kk < p.nkk && re < p.nRe) ? p.phE[size_t(kk) * p.nRe + re] : make_double2(0., 0.);
        const double v = (m < NKK) ? (s ? -P.y : P.x) : (s ? P.x : P.y);
        As[m * LDA + k] = T(v);
      }
      __syncthreads();
      if (ch + 1 < nchunk) load_b(bn, re0 + 8);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        T a[MTc];
#pragma unroll
        for (int m = 0; m < MTc; ++m) a[m] = As[(16 * m + lr) * LDA + 4 * ks + lg];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const T b = (ks & 1) ? bq[ks >> 1][j].y : bq[ks >> 1][j].x;
#pragma unroll
          for (int m = 0; m < MTc; ++m) acc[m][j] = MfmaC<T>::mac(a[m], b, acc[m][j]);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j) bq[h][j] = bn[h][j];
    }
    // epilogue in passes of 16 outer points: real tile q and imaginary tile MTc/2 + q
#pragma unroll
    for (int q = 0; q < MTc / 2; ++q) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int cl = (wave * 4 + j) * 16 + lr;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = MfmaC<T>::row(lg, i);
          Cb[r * 256 + cl] = acc[q][j][i];
          Cb[(16 + r) * 256 + cl] = acc[MTc / 2 + q][j][i];
        }
      }
      __syncthreads();
      cache_panel_epilogue<T, OutT>(p, Cb, c0, 16 * q, rp, nu, p.nkk);
    }
  }
}

inline size_t cache_lds_bytes(int elem) {
  return std::max(size_t(2 * kCacheNKK) * 17 * elem, size_t(2) * 16 * 256 * elem);
}

template <typename T, typename OutT>
void launch_cache(const CacheParams& p, hipStream_t s) {
  PHX_REQUIRE(p.nw <= 32, "cacheElPh kernel supports nw <= 32");
  PHX_REQUIRE(p.nkk <= kCacheNKK, "cacheElPh kernel: too many outer points per launch");
  if (p.nkk <= 0) return;
  const size_t lds = cache_lds_bytes(sizeof(T));
  auto kern = k_cache<T, OutT>;
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(kern, dim3(p.nRp * p.nm), dim3(256), lds, s, p);
  PHX_HIP(hipGetLastError());
}
template void launch_cache<double, double>(const CacheParams&, hipStream_t);
template void launch_cache<float, float>(const CacheParams&, hipStream_t);

// ------------------------------------------------------------------ cacheElPh on mesh lines
// Outer points that share the mesh line (j1, j2) and differ in j0 factor the
// phase: with n_d = b_d.R_e / 2pi,
//   e^{ik1.R_e} = e^{2 pi i j0 n0/m0} e^{2 pi i (j1 n1/m1 + j2 n2/m2)}.
// Stage a (one HBM stream over gW for up to kLinesPerBatch = 16 lines):
//   Tg(l, g, c) = sum_{R_e in group g} e^{2 pi i (j1 n1/m1 + j2 n2/m2)}/de gW(R_e, c)
// with groups of equal n0 mod m0 (about 2 * 10 + 1 of them for a 10^3 R_e grid).
// Stage b (per outer point): g1(k1, c) = sum_g e^{2 pi i j0 n0(g)/m0} Tg(l(k1), g, c),
// then the same U1 / s(k1, Rp) panel epilogue as k_cache.
constexpr int kLineDepth = 4;  // gW rows in flight per thread (16 bytes each)

// NL lines, CPT adjacent columns per thread (one 16-byte load per R_e for
// CPT * sizeof(T2) == 16).  The phase table is (R_e, line): the NL factors of
// one R_e are contiguous, read by the whole workgroup alike.
template <typename T, int NL, int CPT>
__global__ void __launch_bounds__(256) k_cache_line_a(CacheParams p) {
  typedef typename Vec2<T>::type T2;
  typedef T vec __attribute__((ext_vector_type(2 * CPT)));
  constexpr int D = kLineDepth;
  const size_t col0 = (size_t(blockIdx.x) * 256 + threadIdx.x) * CPT;
  const bool ok = col0 < p.N;  // N is a multiple of CPT
  const T2* G = static_cast<const T2*>(p.gW);
  const T2* ph = static_cast<const T2*>(p.phLine);
  T2* out = static_cast<T2*>(p.Tg);
  vec cur[D], nxt[D];
  auto load = [&](vec(&b)[D], int i0) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int i = i0 + u;
      if (ok && i < p.nRe) b[u] = *reinterpret_cast<const vec*>(G + size_t(p.reOrder[i]) * p.N + col0);
      else b[u] = vec(T(0));
    }
  };
  load(cur, 0);
  T ar[NL][CPT], ai[NL][CPT];
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int c = 0; c < CPT; ++c) ar[l][c] = ai[l][c] = T(0);
  for (int i0 = 0; i0 < p.nRe; i0 += D) {
    load(nxt, i0 + D);
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int i = i0 + u;
      if (i < p.nRe) {
#pragma unroll
        for (int l = 0; l < NL; ++l) {
          const T2 f = ph[size_t(i) * NL + l];
#pragma unroll
          for (int c = 0; c < CPT; ++c) {
            ar[l][c] += f.x * cur[u][2 * c] - f.y * cur[u][2 * c + 1];
            ai[l][c] += f.x * cur[u][2 * c + 1] + f.y * cur[u][2 * c];
          }
        }
        const int g = p.reGroupEnd[i];
        if (g >= 0) {
#pragma unroll
          for (int l = 0; l < NL; ++l) {
            if (ok) {
              vec o;
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                o[2 * c] = ar[l][c];
                o[2 * c + 1] = ai[l][c];
              }
              *reinterpret_cast<vec*>(out + (size_t(l) * p.nGe + g) * p.N + col0) = o;
            }
#pragma unroll
            for (int c = 0; c < CPT; ++c) ar[l][c] = ai[l][c] = T(0);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < D; ++u) cur[u] = nxt[u];
  }
}

// Stage a for 16 lines on the matrix cores (FP32 data).  Per R_e group g,
// Tg(l, g, c) = sum_{R_e in g} ph(R_e, l) gW(R_e, c) is a complex GEMM with
// M = 16 lines, K = the group's R_e and N = the gW columns.  It runs as real
// products on v_mfma_f32_16x16x4f32, four R_e at a time (groups padded to a
// multiple of 4 with zero entries): lane (n, kq) = (L % 16, L / 16) loads
// gW(R_kq, column n) of each 16-column tile (8 bytes, both components) and
// ph(R_kq, line n), and two K-steps take the x and the y components,
//   Re += a x - b y,   Im += b x + a y      (ph = a + i b, gW = x + i y),
// with A = (a | -b) and (b | a).  D(line 4 kq + i, column n) holds Tg.  A wave
// owns kLineMT tiles of 16 columns and keeps the next block's loads in flight;
// the padded order sits in LDS.  8 bytes of gW per element against 128 FLOPs
// on the matrix cores: bound by HBM.
constexpr int kLineMT = 8;

__global__ void __launch_bounds__(256, 4) k_cache_line_am(CacheParams p) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char smem[];
  int2* sPad = reinterpret_cast<int2*>(smem);  // nRePad
  for (int i = threadIdx.x; i < p.nRePad; i += 256) sPad[i] = p.rePad[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, n = lane & 15, kq = lane >> 4;
  const size_t colw = (size_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * (16 * kLineMT);
  if (colw >= p.N) return;  // whole waves
  const float2* G = static_cast<const float2*>(p.gW);
  const float2* ph = static_cast<const float2*>(p.phLine);
  float2* out = static_cast<float2*>(p.Tg);
  const int nBlk = p.nRePad / 4;
  auto load = [&](int blk, float2(&v)[kLineMT], float2& f) {
    const int2 e = blk < nBlk ? sPad[4 * blk + kq] : make_int2(-1, -1);
#pragma unroll
    for (int t = 0; t < kLineMT; ++t) {
      const size_t c = colw + 16 * t + n;
      v[t] = (e.x >= 0 && c < p.N) ? G[size_t(e.x) * p.N + c] : make_float2(0.f, 0.f);
    }
    f = e.y >= 0 ? ph[size_t(e.y) * 16 + n] : make_float2(0.f, 0.f);
  };
  f4 re[kLineMT], im[kLineMT];
#pragma unroll
  for (int t = 0; t < kLineMT; ++t) re[t] = im[t] = f4{0.f, 0.f, 0.f, 0.f};
  float2 x[kLineMT], xn[kLineMT], f, fn;
  load(0, x, f);
  for (int b = 0; b < nBlk; ++b) {
    load(b + 1, xn, fn);
#pragma unroll
    for (int t = 0; t < kLineMT; ++t) {
      re[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.x, x[t].x, re[t], 0, 0, 0);
      re[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(-f.y, x[t].y, re[t], 0, 0, 0);
      im[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.y, x[t].x, im[t], 0, 0, 0);
      im[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.x, x[t].y, im[t], 0, 0, 0);
    }
    const int g = p.reBlkEnd[b];
    if (g >= 0) {
#pragma unroll
      for (int t = 0; t < kLineMT; ++t) {
        const size_t c = colw + 16 * t + n;
        if (c < p.N) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            out[(size_t(4 * kq + i) * p.nGe + g) * p.N + c] = make_float2(re[t][i], im[t][i]);
        }
        re[t] = im[t] = f4{0.f, 0.f, 0.f, 0.f};
      }
    }
#pragma unroll
    for (int t = 0; t < kLineMT; ++t) x[t] = xn[t];
    f = fn;
  }
}

// One workgroup per (Rp, nu) block of nw^2 columns, in passes of 256 columns
// and, per line, of 16 outer points.  With at most kLineGe R_e groups, a
// thread's partial sums of the line (one per group) are loaded once, all at
// once, into registers and serve every batch of 16 outer points; the batch's
// factors e^{2 pi i j0 n0 / m0} are staged in LDS.
constexpr int kLineGe = 32;

template <typename T, typename OutT>
__global__ void __launch_bounds__(256) k_cache_line_b(CacheParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2;
  const int rp = blockIdx.x / p.nm, nu = blockIdx.x % p.nm;
  const int E = p.nw * p.nw, t = threadIdx.x;
  T* Cb = reinterpret_cast<T*>(smem);                                        // 2 x 16 x 256
  T2* sw = reinterpret_cast<T2*>(smem + size_t(2) * 16 * 256 * sizeof(T));  // 16 x kLineGe
  const T2* Tg = static_cast<const T2*>(p.Tg);
  const bool inReg = p.nGe <= kLineGe;
  for (int c0 = 0; c0 < E; c0 += cache_pass_cols(p.nw)) {
    const int cl = c0 + t;
    const bool ok = cl < E;
    const size_t colg = (size_t(rp) * p.nm + nu) * E + cl;
    for (int l = 0; l < p.nLinesB; ++l) {
      const int kend = p.lineKK[l + 1];
      T2 v[kLineGe];
      if (inReg) {
#pragma unroll
        for (int g = 0; g < kLineGe; ++g)
          v[g] = (ok && g < p.nGe) ? Tg[(size_t(l) * p.nGe + g) * p.N + colg] : T2{T(0), T(0)};
      }
      for (int kbase = p.lineKK[l]; kbase < kend; kbase += 16) {
        T accr[16], acci[16];
#pragma unroll
        for (int kl = 0; kl < 16; ++kl) accr[kl] = acci[kl] = T(0);
        if (inReg) {
          __syncthreads();  // the previous batch's readers of sw are done
          for (int i = t; i < 16 * p.nGe; i += 256) {
            const int kl = i / p.nGe, g = i % p.nGe;
            const double2 w = p.wLine[size_t(min(kbase + kl, kend - 1)) * p.nGe + g];
            sw[kl * kLineGe + g] = T2{T(w.x), T(w.y)};
          }
          __syncthreads();
#pragma unroll
          for (int g = 0; g < kLineGe; ++g) {
            if (g < p.nGe) {
#pragma unroll
              for (int kl = 0; kl < 16; ++kl) {
                const T2 w = sw[kl * kLineGe + g];
                accr[kl] += w.x * v[g].x - w.y * v[g].y;
                acci[kl] += w.x * v[g].y + w.y * v[g].x;
              }
            }
          }
        } else {
          for (int g = 0; g < p.nGe; ++g) {
            const T2 vg = ok ? Tg[(size_t(l) * p.nGe + g) * p.N + colg] : T2{T(0), T(0)};
#pragma unroll
            for (int kl = 0; kl < 16; ++kl) {
              const int kk = min(kbase + kl, kend - 1);
              const double2 w = p.wLine[size_t(kk) * p.nGe + g];
              const T wr = T(w.x), wi = T(w.y);
              accr[kl] += wr * vg.x - wi * vg.y;
              acci[kl] += wr * vg.y + wi * vg.x;
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int kl = 0; kl < 16; ++kl) {
          Cb[kl * 256 + t] = accr[kl];
          Cb[(16 + kl) * 256 + t] = acci[kl];
        }
        __syncthreads();
        cache_panel_epilogue<T, OutT>(p, Cb, c0, kbase, rp, nu, kend);
      }
    }
  }
}

template <typename T, int NL>
static void launch_line_a(const CacheParams& p, hipStream_t s) {
  if constexpr (sizeof(T) == 4 && NL == 16) {
    if (p.rePad && p.reBlkEnd && p.nRePad % 4 == 0) {
      const size_t lds = size_t(p.nRePad) * sizeof(int2);
      PHX_REQUIRE(lds <= 160 * 1024, "cacheElPh line: too many R_e for the matrix-core stage");
      PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cache_line_am),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(std::max<size_t>(lds, 16))));
      hipLaunchKernelGGL(k_cache_line_am, dim3(ceil_div((long long)p.N, 4 * 16 * kLineMT)), dim3(256), lds, s,
                         p);
      PHX_HIP(hipGetLastError());
      return;
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (p.N % 2 == 0) {
      hipLaunchKernelGGL((k_cache_line_a<T, NL, 2>), dim3(ceil_div((long long)p.N / 2, 256)), dim3(256),
                         0, s, p);
      PHX_HIP(hipGetLastError());
      return;
    }
  }
  {
    hipLaunchKernelGGL((k_cache_line_a<T, NL, 1>), dim3(ceil_div((long long)p.N, 256)), dim3(256), 0, s, p);
  }
  PHX_HIP(hipGetLastError());
}

template <typename T, typename OutT>
void launch_cache_line(const CacheParams& p, hipStream_t s) {
  PHX_REQUIRE(p.nw <= 32, "cacheElPh kernel supports nw <= 32");
  PHX_REQUIRE(p.nGe >= 1 && p.N == size_t(p.nRp) * p.nm * p.nw * p.nw, "cacheElPh line: bad sizes");
  PHX_REQUIRE(p.nLinesB >= 1 && p.nLinesB <= kLinesPerBatch && p.nLinesA == line_count_padded(p.nLinesB),
              "cacheElPh line: 1 to 16 lines per batch");
  if (p.nkk <= 0) return;
  switch (p.nLinesA) {
    case 1: launch_line_a<T, 1>(p, s); break;
    case 2: launch_line_a<T, 2>(p, s); break;
    case 3: launch_line_a<T, 3>(p, s); break;
    case 4: launch_line_a<T, 4>(p, s); break;
    case 6: launch_line_a<T, 6>(p, s); break;
    case 8: launch_line_a<T, 8>(p, s); break;
    case 12: launch_line_a<T, 12>(p, s); break;
    default: launch_line_a<T, 16>(p, s); break;
  }
  const size_t lds = size_t(2) * 16 * 256 * sizeof(T) + size_t(16) * kLineGe * 2 * sizeof(T);
  auto kb = k_cache_line_b<T, OutT>;
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kb),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(kb, dim3(p.nRp * p.nm), dim3(256), lds, s, p);
  PHX_HIP(hipGetLastError());
}
template void launch_cache_line<double, double>(const CacheParams&, hipStream_t);
template void launch_cache_line<float, float>(const CacheParams&, hipStream_t);

// FP64 -> FP32 copy of the coupling tensor (the reduced-precision GEMMs read it).
__global__ void k_to_float(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    b[i] = float(a[i]);
}
void launch_to_float(const double* a, float* b, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_float, dim3(8192), dim3(256), 0, s, a, b, n);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ coupling
// (coupling_kernel.hpp, instantiated in coupling_f64.hip and coupling_f32.hip)

// ------------------------------------------------------------------ collect

// One workgroup per (k1, b1): 256 threads stride over the partial slots of the
// panels of b1's degenerate group, then a fixed-order tree sum (deterministic).
__global__ void __launch_bounds__(256) k_collect(CollectParams p) {
  __shared__ double red[4][kCMax];
  const int kk = blockIdx.x, b1 = blockIdx.y;
  if (kk >= p.nkk || b1 >= p.kkNb1[kk]) return;  // uniform per workgroup
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int pan = p.kkPanelBase[kk] + b1;
  const int gs = p.panelGStart[pan], gz = p.panelGSize[pan];
  double sum[kCMax] = {0., 0., 0., 0.};
  const int n = gz * p.nMtiles;
  for (int i = t; i < n; i += 256) {
    const int j = i / p.nMtiles, mt = i - j * p.nMtiles;
    const int src = p.kkPanelBase[kk] + gs + j;
    const double2* q = reinterpret_cast<const double2*>(
        p.partial + (size_t(src) * p.nMtiles + mt) * kGMax * kCMax + (b1 - gs) * kCMax);
    const double2 v0 = q[0], v1 = q[1];
    sum[0] += v0.x;
    sum[1] += v0.y;
    sum[2] += v1.x;
    sum[3] += v1.y;
  }
#pragma unroll
  for (int c = 0; c < kCMax; ++c) {
    double v = sum[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (t < p.nCalc)
    p.linewidth[size_t(t) * p.numStates + p.kkStateOff[kk] + b1] +=
        ((red[0][t] + red[1][t]) + (red[2][t] + red[3][t]));
}

void launch_collect(const CollectParams& p, hipStream_t s) {
  if (p.nkk <= 0) return;
  hipLaunchKernelGGL(k_collect, dim3(p.nkk, p.nb1Max), dim3(256), 0, s, p);
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
