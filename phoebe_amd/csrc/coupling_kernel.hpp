// R_ph Fourier transform + rotations + |g|^2 (+ fused symmetrization and
// Gaussian-delta reduction) as one MFMA kernel template; instantiated per
// GEMM precision in coupling_f64.hip / coupling_f32.hip.
#pragma once

#include <algorithm>

#include "phx_common.hpp"

namespace phx {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

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

template <typename T, int MT, int NTW, bool REDUCE, int NMX>
__global__ void __launch_bounds__(256, 2) k_coupling(CoupParams p) {
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

  // A tile (ROWS x kKC) double-buffered in LDS; B fragments, which belong to
  // one wave only, go straight to registers one K chunk ahead of their use.
  constexpr int LDA = kKC + 1;
  constexpr int KS = kKC / 4;              // MFMA k-steps per chunk
  constexpr int AE = ROWS * kKC / 256;     // A elements loaded per thread per chunk
  T* As0 = reinterpret_cast<T*>(smem);
  T* As1 = As0 + ROWS * LDA;
  const T* A = static_cast<const T*>(p.A) + size_t(row0) * p.lda;
  const T* B = static_cast<const T*>(p.B) + size_t(panel) * p.panelStride;
  const int lr = lane & 15, lk = lane >> 4;
  T areg[AE];
  T b0[KS][NTW], b1[KS][NTW];

  auto load_a = [&](int kc) {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int i = t + 256 * e, r = i / kKC, k = i % kKC;
      areg[e] = A[size_t(r) * p.lda + kc + k];
    }
  };
  auto store_a = [&](T* As) {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int i = t + 256 * e, r = i / kKC, k = i % kKC;
      As[r * LDA + k] = areg[e];
    }
  };
  auto load_b = [&](T(&b)[KS][NTW], int kc) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const T* Bk = B + size_t(kc + ks * 4 + lk) * PWp + lr;
#pragma unroll
      for (int j = 0; j < NTW; ++j)
        if (wave + 4 * j < ntiles) b[ks][j] = Bk[(wave + 4 * j) * 16];
    }
  };
  auto compute = [&](const T* As, const T(&b)[KS][NTW]) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      T a[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) a[m] = As[(16 * m + lr) * LDA + ks * 4 + lk];
#pragma unroll
      for (int j = 0; j < NTW; ++j)
        if (wave + 4 * j < ntiles) {
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[m][j] = Mfma4<T>::mac(a[m], b[ks][j], acc[m][j]);
        }
    }
  };

  const int nch = p.Kp / kKC;
  load_a(0);
  load_b(b0, 0);
  store_a(As0);
  __syncthreads();
  for (int c = 0; c < nch; c += 2) {
    if (c + 1 < nch) {
      load_a((c + 1) * kKC);
      load_b(b1, (c + 1) * kKC);
    }
    compute(As0, b0);
    if (c + 1 < nch) store_a(As1);
    __syncthreads();
    if (c + 1 >= nch) break;
    if (c + 2 < nch) {
      load_a((c + 2) * kKC);
      load_b(b0, (c + 2) * kKC);
    }
    compute(As1, b1);
    if (c + 2 < nch) store_a(As0);
    __syncthreads();
  }

  // ---------------- epilogue, 16 rows per pass: g3 rows into LDS (16 x PWp, type T),
  // rotations, |g|^2, then store (calcCouplingSquared) or symmetrize + Gaussian reduce.
  const int nm = p.nm, nw = p.nw;
  constexpr int TPR = 16;  // threads per row
  const int r = t / TPR, sub = t % TPR;
  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];
  T* g3 = reinterpret_cast<T*>(smem);
  T* g = g3 + size_t(r) * PWp;
  double* gs = reinterpret_cast<double*>(g3 + size_t(16) * PWp) + size_t(r) * nm * p.nb2max;
  const int gstart = REDUCE ? p.panelGStart[panel] : 0;
  const int gsize = REDUCE ? p.panelGSize[panel] : 1;
  const double* e1 = REDUCE ? p.e1pool + p.kkStateOff[kk] : nullptr;
  const int nCalc = p.nCalc;
  double part[kGMax][kCMax];
#pragma unroll
  for (int gi = 0; gi < kGMax; ++gi)
#pragma unroll
    for (int c = 0; c < kCMax; ++c) part[gi][c] = 0.;

#pragma unroll
  for (int h = 0; h < MT; ++h) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
      const int tile = wave + 4 * j;
      if (tile < ntiles) {
#pragma unroll
        for (int i = 0; i < 4; ++i) g3[Mfma4<T>::row(lk, i) * PWp + tile * 16 + lr] = acc[h][j][i];
      }
    }
    __syncthreads();
    const int row = row0 + 16 * h + r;
    const bool valid = row < p.M;
    int iq = 0;
    if (valid) iq = p.qmode ? mesh_diff_index(p.rowMesh[row], p.kkMesh[kk], p.m0, p.m1, p.m2) : p.rowE3[row];
    const double2* e3 = p.E3pool + size_t(iq) * nm * nm;

    // g4(nu2, w2) = sum_nu e3(nu, nu2) g3(nu, w2)  (interaction_elph.cpp:381-390), in place per w2
    if (valid) {
      for (int w2 = sub; w2 < nw; w2 += TPR) {
        double2 x[NMX];
#pragma unroll
        for (int nu = 0; nu < NMX; ++nu)
          if (nu < nm) x[nu] = make_double2(double(g[2 * (nu * nw + w2)]), double(g[2 * (nu * nw + w2) + 1]));
        for (int nu2 = 0; nu2 < nm; ++nu2) {
          double2 s = make_double2(0., 0.);
#pragma unroll
          for (int nu = 0; nu < NMX; ++nu)
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
    if (!REDUCE) continue;
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
  }
  if (!REDUCE) return;
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

inline size_t coupling_lds_bytes(int elem, int MT, int ldb, int nm, int nb2max, bool reduce) {
  const int ROWS = 16 * MT;
  size_t a = 2 * size_t(ROWS) * (kKC + 1) * elem;
  size_t e = size_t(16) * ldb * elem + (reduce ? size_t(16) * nm * nb2max * sizeof(double) : 0);
  size_t r = size_t(kGMax) * kCMax * 4 * sizeof(double);
  return std::max(std::max(a, e), r);
}

template <typename T, int MT, int NTW, bool REDUCE, int NMX>
static void launch_coupling_t(const CoupParams& p, int grid, size_t lds, hipStream_t s) {
  auto kern = k_coupling<T, MT, NTW, REDUCE, NMX>;
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, p);
  PHX_HIP(hipGetLastError());
}

// Panels up to 384 real columns (at most 6 column tiles per wave) use 32-row
// tiles, wider panels 16-row tiles, so that accumulators stay in registers.
template <typename T, bool REDUCE, int NMX>
static void dispatch_ntw(int ntw, const CoupParams& p, int grid, size_t lds, hipStream_t s) {
  switch (ntw) {
#define PHX_NTW(mt, n) case n: launch_coupling_t<T, mt, n, REDUCE, NMX>(p, grid, lds, s); break;
    PHX_NTW(2, 1) PHX_NTW(2, 2) PHX_NTW(2, 3) PHX_NTW(2, 4) PHX_NTW(2, 5) PHX_NTW(2, 6)
    PHX_NTW(1, 7) PHX_NTW(1, 8) PHX_NTW(1, 9)
#undef PHX_NTW
    default: throw Error(-1, "coupling kernel: panel too wide");
  }
}

inline int coupling_rows_per_tile(int ldb) { return ldb <= 384 ? 32 : 16; }

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
  if (p.nm <= 12) {
    if (reduce) dispatch_ntw<T, true, 12>(ntw, p, grid, lds, s);
    else dispatch_ntw<T, false, 12>(ntw, p, grid, lds, s);
  } else if (p.nm <= 18) {
    if (reduce) dispatch_ntw<T, true, 18>(ntw, p, grid, lds, s);
    else dispatch_ntw<T, false, 18>(ntw, p, grid, lds, s);
  } else {
    if (reduce) dispatch_ntw<T, true, 36>(ntw, p, grid, lds, s);
    else dispatch_ntw<T, false, 36>(ntw, p, grid, lds, s);
  }
}


}  // namespace phx
