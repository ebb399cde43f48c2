// Coupling kernels for one batch of outer points k1 (panels = (k1, b1) pairs):
//
//  * k_coupling  : direct R_ph transform, g3 = A B' with A(row, .) = e^{ik2.Rp}
//                  (or e^{iq.Rp} for explicit q) on v_mfma_{f64,f32}_16x16x4,
//                  followed by the shared epilogue;
//  * k_sep_*     : the same g3 for k2 on the Gamma-centred mesh as three short
//                  lattice-direction transforms e^{ik2.Rp} = prod_d w_d^{j_d n_d}
//                  (w_d = e^{2 pi i / m_d}, n = integer coordinates of Rp folded
//                  mod m), the last one fused with the shared epilogue.
//
// Shared epilogue (per 16 rows, g3(nu, w2) of each row in LDS):
//   g'(nu, b2)  = sum_w2 conj(U2(w2, b2)) g3(nu, w2)    (interaction_elph.cpp:393-402)
//   g(nu2, b2)  = sum_nu e3(nu, nu2) g'(nu, b2)         (interaction_elph.cpp:381-390)
// i.e. the reference's two rotations applied in the cheaper order (the electron
// rotation first shrinks w2 -> b2), then |g|^2 and either a store in the
// getCouplingSquared layout or symmetrizeCoupling along b2 and nu plus the
// Gaussian-delta weights of el_scattering.cpp:277-380 into per-workgroup partials.
//
// Instantiated per precision of the transforms in coupling_f64.hip / coupling_f32.hip.
#pragma once

#include <algorithm>
#include <type_traits>

#include "elph_params.hpp"
#include "phx_common.hpp"

namespace phx {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int kNmMax = 36;
constexpr int kKC = 16;   // K chunk staged in LDS

struct CoupParams {
  // direct transform operands
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
  int gsizeMax;                           // largest panelGSize of the launch (0: unknown, kGMax)
  int epRows;                             // k_sep_ep: rows per pass, waves x (rows per wave), rows per wave <= 64 / nm
  double* partial;                        // (L = panel*nMtiles + mtile, kGMax, kCMax)
  int nb2max;
  // separable mesh transform
  const int* meshToRow;                   // full mesh index -> row (active point) or -1
  const void* T2; size_t t2PanelStride;   // (panel, f0 group, jy, jz, col) complex
  int nG0; const int* g0val;              // folded x coordinate of each group
  const double2* tw0;                     // e^{2 pi i t / m0}, t < m0
  int rowLo, rowHi;                       // rows (active indices) taken by the mesh transform
  double screen;                          // energy screen in units of sigma (0: off)
  double* screenOf;                       // per outer state: screen width (sigma) of the panel whose b1 group
                                          // starts there (k_sep_xmin); null: `screen` for every panel
  double* xminPart;                       // (panel, kXminBlocks) closest approach to resonance, energy units
  const double* qWlo; const double* qWhi; // per q: lowest / highest omega >= cutoff
  const double2* rowE2range;              // per row: (lowest, highest) band energy
  // learned screen (k_sep_xm): per (panel, yz) slot masks
  int bf16x3;                             // x-stage GEMMs in split-BF16x3 (FP32 data)
  const unsigned char* slotOn;            // compute only slots with slotOn != 0 (null: all)
  int slotZero;                           // off slots: 1 = write zero partials, 0 = leave untouched
  const int* gate;                        // run only when *gate != 0 (null: always)
  unsigned long long* keptRows;           // rows computed by the last stage are added here (null: off)
  int wideCols;                           // k_coupling_wide: real columns per mode window (multiple of 16)
  // precomputed row lists of the last stage (m0 <= 64)
  unsigned long long* keepMask;           // per slot: bit jx set = row jx kept (null: screen in k_sep_xm)
  const int4* meshRow;                    // per mesh point: (row, nb2, u2off, stateOff), row -1 if inactive
  // adaptive Gaussian smearing: width from the group velocities of the two states
  int adaptive;                           // 0: fixed Gaussian (invSigma, deltaPref)
  const double* stateV;                   // numStates x 3 group velocities
  double qT[9];                           // reciprocal cell, row i divided by mesh(i), row-major
  double minWidth;                        // smallest width (0.1 meV)
  // degenerate phonon modes: per q and mode, first mode of its group | group size << 8
  const int* qModeGrp;
  // separable transform: phonon eigenvectors (q, nu, nu2) in the transform's precision
  const void* E3T;
  // polar term (direct transform epilogue): per e3 index (row of the per-call
  // API, folded q of the builder) nm complex polarCorrectionPart1 values, zero
  // where the reference skips the term; U1 of the outer point: U1pool +
  // U1off[kk], (nw, nb1) column-major.  polarX null: no polar term.
  const double2* polarX;
  const double2* U1pool;
  const int* U1off;
  // split last stage (k_sep_xg + k_sep_ep): g3 rows of the chunk, (slot, jx, col) complex T
  void* G3;
  unsigned long long* keptBands;          // k_sep_ep: sum of nb2 over the rows it computed (null: off)
  int epFirst;                            // split epilogue: 1 = k_sep_ep, 0 = k_sep_ep2 where it applies
  int epRot;                              // k_sep_xg also does the electron rotation (g' to HBM, not g3)
  int fuseYX;                             // the y-stage runs inside the x-stage kernel (k_sep_yx; no T2 in HBM)
  unsigned long long* keptSlots;          // k_sep_xg: slots (panel, yz line) it computed (null: off)
};

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

// Energy-screen width, in units of sigma, of the panel whose b1 group starts
// at outer state is1 (see k_sep_xmin).
__device__ inline double screen_width(const CoupParams& p, int is1) {
  return p.screenOf ? p.screenOf[is1] : p.screen;
}

__device__ inline double gaussian_delta(double energy, double invW, double pref) {
  // GaussianDeltaFunction::getSmearing, delta_function.cpp:40-45
  double x = energy * invW;
  if (x > 6.) return 0.;
  return pref * exp(-x * x);
}

// AdaptiveGaussianDeltaFunction::getSmearing (delta_function.cpp:66-97) with
// v = v(state 1) - v(state 2) (el_scattering.cpp:291-295): width
// sqrt(sum_i (qT_i . v)^2 / 6), at least minWidth; cut at 2 widths and
// renormalized by erf(2).
struct AdaptiveWidth {
  double sigma;  // 0: delta is zero; < 0: zero velocity difference
};
__device__ inline AdaptiveWidth adaptive_width(const CoupParams& p, int is1, int is2) {
  const double v0 = p.stateV[3 * is1] - p.stateV[3 * is2];
  const double v1 = p.stateV[3 * is1 + 1] - p.stateV[3 * is2 + 1];
  const double v2 = p.stateV[3 * is1 + 2] - p.stateV[3 * is2 + 2];
  AdaptiveWidth w;
  if (v0 == 0. && v1 == 0. && v2 == 0.) {
    w.sigma = -1.;
    return w;
  }
  double s = 0.;
  for (int i = 0; i < 3; ++i) {
    const double d = p.qT[3 * i] * v0 + p.qT[3 * i + 1] * v1 + p.qT[3 * i + 2] * v2;
    s += d * d;
  }
  s = sqrt(s / 6.);
  if (s != 0. && s < p.minWidth) s = p.minWidth;
  w.sigma = s;
  return w;
}
__device__ inline double adaptive_delta(double energy, AdaptiveWidth w) {
  if (w.sigma < 0.) return energy == 0. ? 1. : 0.;
  if (w.sigma == 0.) return 0.;
  if (fabs(energy) > 2. * w.sigma) return 0.;
  const double x = energy / w.sigma;
  if (x > 6.) return 0.;
  return exp(-x * x) / 1.7724538509055160273 / w.sigma / 0.9953222650189527;
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

// LDS bytes of the shared epilogue (g3 tile, g', |g|^2 per row).
inline size_t epilogue_lds_bytes(int elem, int ldb, int nm, int nb2max) {
  return size_t(16) * ldb * elem + size_t(16) * nm * nb2max * (sizeof(double2) + sizeof(double));
}

// Electron rotation of the modes nuLo .. nuHi-1 for the 16 rows whose g3
// blocks are in LDS (row r at g3 + r*PWp, (re, im) of column (nu, w2) at
// 2((nu - nuLo)*nw + w2)) into g'(nu, b2) in scratch.  Thread t serves row t/16.
template <typename T>
__device__ void epilogue16_u2(const CoupParams& p, const T* g3, int PWp, unsigned char* scratch,
                              int myRow, int nuLo, int nuHi) {
  const int t = threadIdx.x, r = t >> 4, sub = t & 15;
  const int nw = p.nw;
  const int stride = p.nm * p.nb2max;
  double2* gp = reinterpret_cast<double2*>(scratch) + size_t(r) * stride;
  const bool valid = myRow >= 0;
  const int nb2 = valid ? p.rowNb2[myRow] : 0;
  const T* g = g3 + size_t(r) * PWp;
  const double2* U2 = p.U2pool + (valid ? p.rowU2off[myRow] : 0);
  for (int o = sub; o < (nuHi - nuLo) * nb2; o += 16) {
    const int nl = o / nb2, b2 = o % nb2;
    double2 s = make_double2(0., 0.);
    for (int w2 = 0; w2 < nw; ++w2) {
      const double2 u = U2[w2 + nw * b2];
      const double xr = double(g[2 * (nl * nw + w2)]), xi = double(g[2 * (nl * nw + w2) + 1]);
      s.x += u.x * xr + u.y * xi;
      s.y += u.x * xi - u.y * xr;
    }
    gp[(nuLo + nl) * nb2 + b2] = s;
  }
}

// Rest of the epilogue once g'(nu, b2) of every mode is in scratch: phonon
// rotation, |g|^2, then the store or the symmetrization and the weights.
// Must be called by all 256 threads.
template <bool REDUCE>
__device__ void epilogue16_tail(const CoupParams& p, unsigned char* scratch, int panel, int kk, int b1p,
                                int myRow, int myIq, double (&part)[kGMax][kCMax]) {
  const int t = threadIdx.x, r = t >> 4, sub = t & 15;
  const int nm = p.nm;
  const int stride = nm * p.nb2max;
  double2* gp = reinterpret_cast<double2*>(scratch) + size_t(r) * stride;
  double* gs = reinterpret_cast<double*>(reinterpret_cast<double2*>(scratch) + size_t(16) * stride) +
               size_t(r) * stride;
  const bool valid = myRow >= 0;
  const int nb2 = valid ? p.rowNb2[myRow] : 0;
  const double2* e3 = p.E3pool + size_t(myIq) * nm * nm;
  const bool polar = p.polarX != nullptr && valid;
  for (int o = sub; o < nm * nb2; o += 16) {
    const int nu2 = o / nb2, b2 = o % nb2;
    double2 s = make_double2(0., 0.);
    for (int nu = 0; nu < nm; ++nu) {
      const double2 e = e3[nu + nm * nu2], x = gp[nu * nb2 + b2];
      s.x += x.x * e.x - x.y * e.y;
      s.y += x.x * e.y + x.y * e.x;
    }
    if (polar) {
      // gFinal += x(nu2) <U2(b2)|U1(b1)> (interaction_elph.cpp:406-413 with
      // polarCorrectionPart2, :205-227)
      const double2* u1 = p.U1pool + p.U1off[kk] + size_t(p.nw) * b1p;
      const double2* u2 = p.U2pool + p.rowU2off[myRow] + size_t(p.nw) * b2;
      double2 ov = make_double2(0., 0.);
      for (int w = 0; w < p.nw; ++w) {
        const double2 a = u2[w], b = u1[w];
        ov.x += a.x * b.x + a.y * b.y;
        ov.y += a.x * b.y - a.y * b.x;
      }
      const double2 x = p.polarX[size_t(myIq) * nm + nu2];
      s.x += x.x * ov.x - x.y * ov.y;
      s.y += x.x * ov.y + x.y * ov.x;
    }
    const double v = s.x * s.x + s.y * s.y;
    if (REDUCE) {
      gs[nu2 * nb2 + b2] = v;
    } else {
      const int nb1 = p.kkNb1[kk];
      p.out[p.rowOutOff[myRow] + b1p + size_t(nb1) * (b2 + size_t(nb2) * nu2)] = v;
    }
  }
  if (!REDUCE) return;
  __syncthreads();

  const double* w3 = p.Wpool + size_t(myIq) * nm;
  const double* e2 = p.e2pool + (valid ? p.rowStateOff[myRow] : 0);
  // symmetrizeCoupling (scattering.cpp:1209-1318) along b2, then nu
  for (int nu = sub; nu < nm && valid; nu += 16) {
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
  for (int b2 = sub; b2 < nb2; b2 += 16) {
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
  const double inv_g = 1. / double(float(gsize));
  for (int o = sub; o < nm * nb2; o += 16) {
    const int nu = o / nb2, b2 = o % nb2;
    const double en3 = w3[nu];
    if (en3 < p.cutoff) continue;
    const double cpl = gs[nu * nb2 + b2] * inv_g;
    const double en2 = e2[b2];
    const int is2 = p.rowStateOff[myRow] + b2;
#pragma unroll
    for (int gi = 0; gi < kGMax; ++gi) {
      if (gi >= gsize) break;
      const double en1 = e1[gstart + gi];
      double d1, d2;
      if (p.adaptive) {
        const AdaptiveWidth aw = adaptive_width(p, p.kkStateOff[kk] + gstart + gi, is2);
        d1 = adaptive_delta(en1 - en2 + en3, aw);
        d2 = adaptive_delta(en1 - en2 - en3, aw);
      } else {
        d1 = gaussian_delta(en1 - en2 + en3, p.invSigma, p.deltaPref);
        d2 = gaussian_delta(en1 - en2 - en3, p.invSigma, p.deltaPref);
      }
      if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
      for (int c = 0; c < kCMax; ++c) {
        if (c >= nCalc) break;
        const double f2 = p.fermi[size_t(c) * p.numStates + is2];
        const double n3 = p.bose[(size_t(myIq) * nm + nu) * nCalc + c];
        part[gi][c] += cpl * ((f2 + n3) * d1 + (1. - f2 + n3) * d2) * p.norm / en3 * 3.14159265358979323846;
      }
    }
  }
}

// Epilogue for the 16 rows whose full g3 rows are in LDS (row r at g3 + r*PWp,
// (re, im) of column (nu, w2) at 2(nu*nw + w2)).  Thread t serves row t/16;
// myRow < 0 marks an empty slot.  Must be called by all 256 threads.
template <typename T, bool REDUCE>
__device__ void epilogue16(const CoupParams& p, const T* g3, int PWp, unsigned char* scratch,
                           int panel, int kk, int b1p, int myRow, int myIq,
                           double (&part)[kGMax][kCMax]) {
  epilogue16_u2<T>(p, g3, PWp, scratch, myRow, 0, p.nm);
  __syncthreads();
  epilogue16_tail<REDUCE>(p, scratch, panel, kk, b1p, myRow, myIq, part);
}

// Block reduction of the per-thread partial linewidths into partial[L]
// (CM calculations and GM group members held per thread; the slots beyond
// them are written as zero).
template <int CM, int GM = kGMax, int NWV = 4>
__device__ inline void reduce_partials(const CoupParams& p, int L, unsigned char* smem,
                                       double (&part)[GM][CM]) {
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  double* red = reinterpret_cast<double*>(smem);
  __syncthreads();
#pragma unroll
  for (int gi = 0; gi < GM; ++gi)
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      double v = part[gi][c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[(gi * kCMax + c) * NWV + wave] = v;
    }
  __syncthreads();
  if (t < kGMax * kCMax) {
    const int c = t % kCMax, gi = t / kCMax;
    double v = 0.;
    if (c < CM && gi < GM)
#pragma unroll
      for (int w = 0; w < NWV; ++w) v += red[t * NWV + w];
    p.partial[size_t(L) * kGMax * kCMax + t] = v;
  }
}

// ------------------------------------------------------------------ direct transform
template <typename T, int MT, int NTW, bool REDUCE>
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

  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];
  T* g3 = reinterpret_cast<T*>(smem);
  unsigned char* scratch = smem + size_t(16) * PWp * sizeof(T);
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
    const int row = row0 + 16 * h + (t >> 4);
    const int myRow = row < p.M ? row : -1;
    int myIq = 0;
    if (myRow >= 0)
      myIq = p.qmode ? mesh_diff_index(p.rowMesh[myRow], p.kkMesh[kk], p.m0, p.m1, p.m2) : p.rowE3[myRow];
    epilogue16<T, REDUCE>(p, g3, PWp, scratch, panel, kk, b1p, myRow, myIq, part);
  }
  if (REDUCE) reduce_partials<kCMax>(p, L, smem, part);
}

// Wide panels (2 nm nw > 576): 16-row tiles, the panel columns in windows of
// whole modes (<= 384 real columns), each window's MFMA product rotated by U2
// into g'(nu, b2) before the next; the phonon rotation and the weights run once
// all modes are in.
constexpr int kWideNtw = 6;
template <typename T, bool REDUCE>
__global__ void __launch_bounds__(256, 1) k_coupling_wide(CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NTW = kWideNtw;
  const int WCOL = p.wideCols;  // <= 4 * NTW * 16
  const int nTilesTot = p.nPanels * p.nMtiles;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int L = (bid % 8) * perXcd + bid / 8;
  if (L >= nTilesTot) return;
  const int panel = L / p.nMtiles, mtile = L % p.nMtiles;
  const int row0 = mtile * 16;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15, lk = lane >> 4;
  typedef typename Mfma4<T>::acc_t acc_t;
  constexpr int LDA = kKC + 1, KS = kKC / 4, AE = 16 * kKC / 256;
  T* As = reinterpret_cast<T*>(smem);                                    // 16 x LDA
  T* g3 = As + 16 * LDA;                                                 // 16 x WCOL
  unsigned char* scratch = reinterpret_cast<unsigned char*>(g3 + 16 * WCOL);
  scratch = smem + (size_t(scratch - smem) + 15) / 16 * 16;
  const T* A = static_cast<const T*>(p.A) + size_t(row0) * p.lda;
  const T* B = static_cast<const T*>(p.B) + size_t(panel) * p.panelStride;
  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];
  const int row = row0 + (t >> 4);
  const int myRow = row < p.M ? row : -1;
  const int nuPerWin = max(1, (WCOL / 2) / p.nw);  // WCOL >= 2 nw
  for (int nu0 = 0; nu0 < p.nm; nu0 += nuPerWin) {
    const int nu1 = min(p.nm, nu0 + nuPerWin);
    const int c0 = 2 * nu0 * p.nw, width = 2 * (nu1 - nu0) * p.nw;
    const int ntiles = (width + 15) / 16;
    acc_t acc[NTW];
#pragma unroll
    for (int j = 0; j < NTW; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[j][i] = 0;
    for (int kc = 0; kc < p.Kp; kc += kKC) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < AE; ++e) {
        const int i = t + 256 * e, r = i / kKC, k = i % kKC;
        As[r * LDA + k] = A[size_t(r) * p.lda + kc + k];
      }
      T b[KS][NTW];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const T* Bk = B + size_t(kc + ks * 4 + lk) * p.ldb + c0 + lr;
#pragma unroll
        for (int j = 0; j < NTW; ++j) {
          const int col = (wave + 4 * j) * 16 + lr;
          b[ks][j] = (wave + 4 * j < ntiles && col < width) ? Bk[(wave + 4 * j) * 16] : T(0);
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const T a = As[lr * LDA + ks * 4 + lk];
#pragma unroll
        for (int j = 0; j < NTW; ++j)
          if (wave + 4 * j < ntiles) acc[j] = Mfma4<T>::mac(a, b[ks][j], acc[j]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
      const int tile = wave + 4 * j;
      if (tile < ntiles)
#pragma unroll
        for (int i = 0; i < 4; ++i) g3[Mfma4<T>::row(lk, i) * WCOL + tile * 16 + lr] = acc[j][i];
    }
    __syncthreads();
    epilogue16_u2<T>(p, g3, WCOL, scratch, myRow, nu0, nu1);
  }
  __syncthreads();
  int myIq = 0;
  if (myRow >= 0)
    myIq = p.qmode ? mesh_diff_index(p.rowMesh[myRow], p.kkMesh[kk], p.m0, p.m1, p.m2) : p.rowE3[myRow];
  double part[kGMax][kCMax];
#pragma unroll
  for (int gi = 0; gi < kGMax; ++gi)
#pragma unroll
    for (int c = 0; c < kCMax; ++c) part[gi][c] = 0.;
  epilogue16_tail<REDUCE>(p, scratch, panel, kk, b1p, myRow, myIq, part);
  if (REDUCE) reduce_partials<kCMax>(p, L, smem, part);
}

inline size_t coupling_wide_lds_bytes(int elem, int nm, int nb2max, int wcol) {
  const size_t a = (size_t(16) * (kKC + 1) + size_t(16) * wcol) * elem;
  return std::max((a + 15) / 16 * 16 + size_t(16) * nm * nb2max * (sizeof(double2) + sizeof(double)),
                  size_t(kGMax) * kCMax * 4 * sizeof(double));
}

inline size_t coupling_lds_bytes(int elem, int MT, int ldb, int nm, int nb2max, bool reduce) {
  (void)reduce;
  const int ROWS = 16 * MT;
  size_t a = 2 * size_t(ROWS) * (kKC + 1) * elem;
  size_t e = epilogue_lds_bytes(elem, ldb, nm, nb2max);
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

// Panels up to 384 real columns (at most 6 column tiles per wave) use 32-row
// tiles, wider panels 16-row tiles, so that accumulators stay in registers.
template <typename T, bool REDUCE>
static void dispatch_ntw(int ntw, const CoupParams& p, int grid, size_t lds, hipStream_t s) {
  switch (ntw) {
#define PHX_NTW(mt, n) case n: launch_coupling_t<T, mt, n, REDUCE>(p, grid, lds, s); break;
    PHX_NTW(2, 1) PHX_NTW(2, 2) PHX_NTW(2, 3) PHX_NTW(2, 4) PHX_NTW(2, 5) PHX_NTW(2, 6)
    PHX_NTW(1, 7) PHX_NTW(1, 8) PHX_NTW(1, 9)
#undef PHX_NTW
    default: throw Error(-1, "coupling kernel: panel too wide");
  }
}

inline int coupling_rows_per_tile(int ldb) { return ldb <= 384 ? 32 : 16; }  // wide panels: 16

template <typename T>
void launch_coupling(const CoupParams& p, bool reduce, hipStream_t s) {
  PHX_REQUIRE(p.nm <= kNmMax, "coupling kernel supports nm <= 36");
  PHX_REQUIRE(p.ldb % 16 == 0 && p.Kp % kKC == 0, "coupling kernel: bad padding");
  if (p.ldb > 9 * 64) {  // wider than one pass of the tiled kernel
    PHX_REQUIRE(p.nMtiles * 16 >= p.M, "coupling kernel: wide panels use 16-row tiles");
    const int tiles = p.nPanels * p.nMtiles;
    if (tiles <= 0) return;
    // widest mode window (whole modes, <= 384 real columns) that fits the LDS budget
    CoupParams q = p;
    q.wideCols = 0;
    size_t lds = 0;
    for (int nuw = std::max(1, 192 / p.nw); nuw >= 1; --nuw) {
      q.wideCols = int(round_up(size_t(2) * nuw * p.nw, 16));
      lds = coupling_wide_lds_bytes(sizeof(T), p.nm, p.nb2max, q.wideCols);
      if (lds <= 160 * 1024) break;
    }
    PHX_REQUIRE(lds <= 160 * 1024, "coupling kernel: nm x bands per inner point too large for LDS");
    auto go = [&](auto kern) {
      PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
      hipLaunchKernelGGL(kern, dim3(ceil_div(tiles, 8) * 8), dim3(256), lds, s, q);
    };
    if (reduce) go(k_coupling_wide<T, true>);
    else go(k_coupling_wide<T, false>);
    PHX_HIP(hipGetLastError());
    return;
  }
  const int MT = coupling_rows_per_tile(p.ldb) / 16;
  const int ntw = ceil_div(p.ldb / 16, 4);
  const int tiles = p.nPanels * p.nMtiles;
  if (tiles <= 0) return;
  const int grid = ceil_div(tiles, 8) * 8;
  const size_t lds = coupling_lds_bytes(sizeof(T), MT, p.ldb, p.nm, p.nb2max, reduce);
  PHX_REQUIRE(lds <= 160 * 1024, "coupling kernel: LDS budget exceeded");
  if (reduce) dispatch_ntw<T, true>(ntw, p, grid, lds, s);
  else dispatch_ntw<T, false>(ntw, p, grid, lds, s);
}

// ------------------------------------------------------------------ separable transform
// Integer layout of the R_ph set folded on the mesh:
//   lines  = distinct (f0, f1), entries (f2, Rp)          (stage z)
//   groups = distinct f0,       entries (f1, line)        (stage y)
struct SepParams {
  const void* B; size_t panelStride; int ldb;   // panels: C(Rp, col) in row 2 Rp
  int ncol, nPanels, mx, my, mz;
  int nLines; const int* lineStart; const int* lineF2; const int* lineRp;
  const double2* twz;                           // e^{2 pi i t / mz}
  void* T1; size_t t1PanelStride;               // (panel, line, jz, col)
  int nG0; const int* grpStart; const int* grpF1; const int* grpLine;
  const double2* twy;                           // e^{2 pi i t / my}
  void* T2; size_t t2PanelStride;               // (panel, group, jy, jz, col)
};

// Register-blocked short transform shared by k_sep_z and k_sep_y:
// out(j, c) = sum_e W(e, j) X(e, c) for j < m, c < ncol, X (the data's
// precision, converted to FP64 on reading) and W in LDS.
// Thread (wave w, lane l) takes the columns c = l, l + 64, ... and the rows
// j = w + 4 i, kSepJB of them at once, so one LDS read of X(e, c) serves
// kSepJB rows and W(e, j) is the same for the whole wave (a broadcast); no
// integer division.  Every output sums its terms in ascending e, as a plain
// loop over the outputs would.
constexpr int kSepJB = 10;
template <typename T, typename Store>
__device__ __forceinline__ void sep_short_transform(const typename Vec2<T>::type* X, const double2* W, int ne,
                                                    int m, int ncol, Store store) {
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  for (int c = l; c < ncol; c += 64) {
    for (int j0 = w; j0 < m; j0 += 4 * kSepJB) {
      double2 acc[kSepJB];
#pragma unroll
      for (int i = 0; i < kSepJB; ++i) acc[i] = make_double2(0., 0.);
      for (int e = 0; e < ne; ++e) {
        const typename Vec2<T>::type xv = X[e * ncol + c];
        const double2 x = make_double2(double(xv.x), double(xv.y));
        const double2* we = W + e * m;
#pragma unroll
        for (int i = 0; i < kSepJB; ++i) {
          const int j = j0 + 4 * i;
          if (j < m) {
            const double2 ww = we[j];
            acc[i].x += ww.x * x.x - ww.y * x.y;
            acc[i].y += ww.x * x.y + ww.y * x.x;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < kSepJB; ++i)
        if (j0 + 4 * i < m) store(j0 + 4 * i, c, acc[i]);
    }
  }
}

// Entries of a short transform into LDS: X(e, c) = row(e)[c] for e < ne,
// c < ncol (the data's precision), and the twiddles W(e, j) = tw[(j f(e)) mod m],
// j < m.  The rows' offsets and the f(e) are staged first (off, fe: ne ints
// each); every thread then keeps up to 8 row loads in flight before its LDS
// stores, so a workgroup waits about one memory latency, not one per entry.
template <typename T2v, typename Meta>
__device__ __forceinline__ void sep_stage_entries(const T2v* base, T2v* X, double2* W, int* off, int* fe, int ne,
                                                  int ncol, int m, const double2* tw, Meta meta) {
  const int t = threadIdx.x, bs = blockDim.x;
  for (int e = t; e < ne; e += bs) meta(e, off[e], fe[e]);
  __syncthreads();
  const int n = ne * ncol;
  for (int i0 = t; i0 < n; i0 += 8 * bs) {
    T2v v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + bs * u;
      if (i < n) {
        const int e = i / ncol;
        v[u] = base[size_t(off[e]) + (i - e * ncol)];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + bs * u;
      if (i < n) X[i] = v[u];
    }
  }
  for (int i = t; i < ne * m; i += bs) {
    const int e = i / m, j = i - e * m;
    W[i] = tw[(j * fe[e]) % m];
  }
  __syncthreads();
}

// LDS of k_sep_z / k_sep_y for ne entries: X (ne x ncol in the data's
// precision), W (ne x m complex FP64), the entries' row offsets and factors.
inline size_t sep_stage_lds_bytes(int elem, int ne, int ncol, int m) {
  const size_t x = (size_t(ne) * ncol * 2 * elem + 15) / 16 * 16;
  return x + size_t(ne) * m * sizeof(double2) + size_t(ne) * 2 * sizeof(int);
}

// T1(line, jz, c) = sum_{(f2, Rp) in line} w_z^{jz f2} C(Rp, c).  The line's
// entries and their twiddles w_z^{jz f2} (an entries x mz table) sit in LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_sep_z(SepParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  const int line = blockIdx.x, panel = blockIdx.y;
  const int e0 = p.lineStart[line], ne = p.lineStart[line + 1] - e0;
  const int ncol = p.ncol, mz = p.mz;
  T2v* X = reinterpret_cast<T2v*>(smem);  // (entries, ncol)
  double2* W = reinterpret_cast<double2*>(smem + (size_t(ne) * ncol * sizeof(T2v) + 15) / 16 * 16);  // (entries, mz)
  int* off = reinterpret_cast<int*>(W + size_t(ne) * mz);
  int* fe = off + ne;
  // C(Rp, c): row 2 Rp of the panel, complex entries (re, im) at 2c, 2c + 1
  const T2v* B = reinterpret_cast<const T2v*>(static_cast<const T*>(p.B) + size_t(panel) * p.panelStride);
  sep_stage_entries(B, X, W, off, fe, ne, ncol, mz, p.twz, [&](int e, int& o, int& f) {
    o = p.lineRp[e0 + e] * p.ldb;  // row 2 Rp: 2 Rp ldb elements of T = Rp ldb complex
    f = p.lineF2[e0 + e];
  });
  T2v* out = reinterpret_cast<T2v*>(static_cast<T*>(p.T1) + size_t(panel) * p.t1PanelStride) +
             size_t(line) * mz * ncol;
  sep_short_transform<T>(X, W, ne, mz, ncol, [&](int jz, int c, double2 v) {
    out[size_t(jz) * ncol + c] = T2v{T(v.x), T(v.y)};
  });
}

// T2(group, jy, jz, c) = sum_{(f1, line) in group} w_y^{jy f1} T1(line, jz, c)
// (entries and twiddles in LDS as in k_sep_z)
template <typename T>
__global__ void __launch_bounds__(256) k_sep_y(SepParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  const int grp = blockIdx.x, jz = blockIdx.y, panel = blockIdx.z;
  const int e0 = p.grpStart[grp], ne = p.grpStart[grp + 1] - e0;
  const int ncol = p.ncol, my = p.my;
  T2v* X = reinterpret_cast<T2v*>(smem);  // (entries, ncol)
  double2* W = reinterpret_cast<double2*>(smem + (size_t(ne) * ncol * sizeof(T2v) + 15) / 16 * 16);  // (entries, my)
  int* off = reinterpret_cast<int*>(W + size_t(ne) * my);
  int* fe = off + ne;
  const T2v* T1 = reinterpret_cast<const T2v*>(static_cast<const T*>(p.T1) + size_t(panel) * p.t1PanelStride);
  sep_stage_entries(T1, X, W, off, fe, ne, ncol, my, p.twy, [&](int e, int& o, int& f) {
    o = (p.grpLine[e0 + e] * p.mz + jz) * ncol;
    f = p.grpF1[e0 + e];
  });
  T2v* out = reinterpret_cast<T2v*>(static_cast<T*>(p.T2) + size_t(panel) * p.t2PanelStride);
  sep_short_transform<T>(X, W, ne, my, ncol, [&](int jy, int c, double2 v) {
    out[((size_t(grp) * my + jy) * p.mz + jz) * ncol + c] = T2v{T(v.x), T(v.y)};
  });
}

// Last stage + epilogue for one (jy, jz) line of mesh points and one panel.
//
//  1. rows of the line: active point, q index, and (screen > 0) the energy
//     screen: a row is kept only if some e1 of the panel's degenerate group,
//     band e2 of the row and mode omega >= cutoff at q give
//     |e1 - e2 +- omega| < screen * sigma (an interval test on the band and
//     mode ranges, so no contributing row is dropped; every skipped term has
//     all its Gaussians below exp(-screen^2) of their peak);
//  2. surviving rows, 8 per pass: g3(jx; c) = sum_groups w_x^{jx f0} T2(group, jy, jz, c);
//  3. epilogue with 32 threads per row: thread nu2 owns the column e3(., nu2) (prefetched),
//     g'(nu, b2) via U2 staged in LDS, g(nu2, b2) and |g|^2 for all b2 in registers,
//     symmetrization along b2 in registers and along nu through LDS (only when the
//     modes at q are degenerate), Gaussian weights.
constexpr int kSepRows = 8;   // rows per pass
constexpr int kSepNb2 = 8;    // largest band count of an inner point on this path
constexpr int kSepNm = 32;    // largest mode count on this path

struct SepXShared {            // per-row data staged in LDS
  int row, iq, nb2, u2off, soff;
};

template <typename T, bool REDUCE, int NMX, int CM>
__global__ void __launch_bounds__(256, 2) k_sep_x(CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  const int nLinesYZ = p.m1 * p.m2;
  const int nTilesTot = p.nPanels * nLinesYZ;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int L = (bid % 8) * perXcd + bid / 8;
  if (L >= nTilesTot) return;
  const int panel = L / nLinesYZ, yz = L % nLinesYZ;
  const int jy = yz % p.m1, jz = yz / p.m1;
  const int t = threadIdx.x;
  const int nm = p.nm, nw = p.nw, ncol = nm * nw;
  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];

  // LDS carve-up
  T2v* X = reinterpret_cast<T2v*>(smem);                        // (groups, ncol)
  T2v* W = X + size_t(p.nG0) * ncol;                             // (m0, groups)
  const size_t xwBytes = (size_t(p.nG0) * ncol + size_t(p.m0) * p.nG0) * sizeof(T2v);
  double2* g3 = reinterpret_cast<double2*>(smem + (xwBytes + 15) / 16 * 16);  // 8 x ncol
  double2* sU2 = g3 + size_t(kSepRows) * ncol;                   // 8 x (nw * nb2max)
  double2* gp = sU2 + size_t(kSepRows) * nw * p.nb2max;          // 8 x (nm * nb2max)
  double* gs = reinterpret_cast<double*>(gp + size_t(kSepRows) * nm * p.nb2max);  // 8 x nm*nb2max
  double* se2 = gs + size_t(kSepRows) * nm * p.nb2max;           // 8 x nb2max
  double* sw3 = se2 + size_t(kSepRows) * p.nb2max;               // 8 x nm
  SepXShared* rinfo = reinterpret_cast<SepXShared*>(sw3 + size_t(kSepRows) * nm);  // m0
  int* list = reinterpret_cast<int*>(rinfo + p.m0);              // m0 + 1 (count at [m0])

  // ---- rows of the line and the screen
  for (int jx = t; jx < p.m0; jx += 256) {
    const int mesh = (jz * p.m1 + jy) * p.m0 + jx;
    int row = p.meshToRow[mesh];
    if (row < p.rowLo || row >= p.rowHi) row = -1;
    SepXShared ri{row, 0, 0, 0, 0};
    if (row >= 0) {
      ri.iq = mesh_diff_index(mesh, p.kkMesh[kk], p.m0, p.m1, p.m2);
      ri.nb2 = p.rowNb2[row];
      ri.u2off = p.rowU2off[row];
      ri.soff = p.rowStateOff[row];
      if (p.screen > 0.) {
        const double wlo = p.qWlo[ri.iq], whi = p.qWhi[ri.iq];
        const double e2lo = p.e2pool[ri.soff], e2hi = p.e2pool[ri.soff + ri.nb2 - 1];
        const double* e1 = p.e1pool + p.kkStateOff[kk] + p.panelGStart[panel];
        const double lim = screen_width(p, int(e1 - p.e1pool)) / p.invSigma;
        bool keep = false;
        if (whi >= wlo) {  // at least one mode above the cutoff
          for (int gi = 0; gi < p.panelGSize[panel]; ++gi) {
            const double a_lo = e1[gi] - e2hi + wlo, a_hi = e1[gi] - e2lo + whi;
            const double b_lo = e1[gi] - e2hi - whi, b_hi = e1[gi] - e2lo - wlo;
            keep |= (a_hi > -lim && a_lo < lim) || (b_hi > -lim && b_lo < lim);
          }
        }
        if (!keep) ri.row = -1;
      }
    }
    rinfo[jx] = ri;
  }
  const T2v* T2g = static_cast<const T2v*>(p.T2) + size_t(panel) * (p.t2PanelStride / 2);
  for (int i = t; i < p.nG0 * ncol; i += 256) {
    const int g = i / ncol, c = i % ncol;
    X[i] = T2g[((size_t(g) * p.m1 + jy) * p.m2 + jz) * ncol + c];
  }
  for (int i = t; i < p.m0 * p.nG0; i += 256) {
    const int jx = i / p.nG0, g = i % p.nG0;
    const double2 w = p.tw0[(jx * p.g0val[g]) % p.m0];
    W[i] = T2v{T(w.x), T(w.y)};
  }
  __syncthreads();
  if (t == 0) {
    int n = 0;
    for (int jx = 0; jx < p.m0; ++jx)
      if (rinfo[jx].row >= 0) list[n++] = jx;
    list[p.m0] = n;
  }
  __syncthreads();
  const int nkeep = list[p.m0];

  double part[kGMax][CM];
#pragma unroll
  for (int gi = 0; gi < kGMax; ++gi)
#pragma unroll
    for (int c = 0; c < CM; ++c) part[gi][c] = 0.;
  const int r = t >> 5, sub = t & 31;   // 32 threads per row in the epilogue
  const int nCalc = p.nCalc;
  const int gstart = p.panelGStart[panel], gsize = p.panelGSize[panel];
  const double* e1 = p.e1pool + p.kkStateOff[kk] + gstart;
  const double inv_g = 1. / double(float(gsize));

  for (int base = 0; base < nkeep; base += kSepRows) {
    // ---- g3 of the pass's rows (16 B of X per 8 complex FMAs)
    for (int c = t; c < ncol; c += 256) {
      double2 acc[kSepRows];
#pragma unroll
      for (int q = 0; q < kSepRows; ++q) acc[q] = make_double2(0., 0.);
      int jxs[kSepRows];
#pragma unroll
      for (int q = 0; q < kSepRows; ++q) jxs[q] = list[min(base + q, nkeep - 1)];
      for (int g = 0; g < p.nG0; ++g) {
        const T2v xv = X[g * ncol + c];
        const double xr = double(xv.x), xi = double(xv.y);
#pragma unroll
        for (int q = 0; q < kSepRows; ++q) {
          const T2v w = W[jxs[q] * p.nG0 + g];
          const double wr = double(w.x), wi = double(w.y);
          acc[q].x += wr * xr - wi * xi;
          acc[q].y += wr * xi + wi * xr;
        }
      }
#pragma unroll
      for (int q = 0; q < kSepRows; ++q) g3[q * ncol + c] = acc[q];
    }
    // ---- per-row prefetch: U2 and energies in LDS, e3 column / omega / Bose in registers
    const bool rowOn = base + r < nkeep;
    SepXShared ri{-1, 0, 0, 0, 0};
    if (rowOn) ri = rinfo[list[base + r]];
    const int nb2 = ri.row >= 0 ? ri.nb2 : 0;
    for (int i = sub; i < nw * nb2; i += 32) sU2[r * nw * p.nb2max + i] = p.U2pool[ri.u2off + i];
    if (sub < nb2) se2[r * p.nb2max + sub] = p.e2pool[ri.soff + sub];
    double2 e3c[NMX];
    double w3 = 0., n3[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) n3[c] = 0.;
    const bool modeOn = nb2 > 0 && sub < nm;
    if (modeOn) {
      const double2* e3 = p.E3pool + size_t(ri.iq) * nm * nm + size_t(nm) * sub;
#pragma unroll
      for (int nu = 0; nu < NMX; ++nu)
        if (nu < nm) e3c[nu] = e3[nu];
      w3 = p.Wpool[size_t(ri.iq) * nm + sub];
      sw3[r * nm + sub] = w3;
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (c < nCalc) n3[c] = p.bose[(size_t(ri.iq) * nm + sub) * nCalc + c];
    }
    __syncthreads();
    // ---- g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2), thread nu = sub
    if (modeOn) {
      const double2* gr = g3 + r * ncol + sub * nw;
      for (int b2 = 0; b2 < nb2; ++b2) {
        const double2* u = sU2 + r * nw * p.nb2max + nw * b2;
        double2 s = make_double2(0., 0.);
        for (int w2 = 0; w2 < nw; ++w2) {
          s.x += u[w2].x * gr[w2].x + u[w2].y * gr[w2].y;
          s.y += u[w2].x * gr[w2].y - u[w2].y * gr[w2].x;
        }
        gp[(r * nm + sub) * p.nb2max + b2] = s;
      }
    }
    __syncthreads();
    // ---- g(nu2, b2) = sum_nu e3(nu, nu2) g'(nu, b2), thread nu2 = sub; |g|^2 in registers
    double v[kSepNb2];
#pragma unroll
    for (int b = 0; b < kSepNb2; ++b) v[b] = 0.;
    if (modeOn) {
#pragma unroll
      for (int b2 = 0; b2 < kSepNb2; ++b2) {
        if (b2 < nb2) {
          double2 s = make_double2(0., 0.);
#pragma unroll
          for (int nu = 0; nu < NMX; ++nu)
            if (nu < nm) {
              const double2 x = gp[(r * nm + nu) * p.nb2max + b2];
              s.x += x.x * e3c[nu].x - x.y * e3c[nu].y;
              s.y += x.x * e3c[nu].y + x.y * e3c[nu].x;
            }
          v[b2] = s.x * s.x + s.y * s.y;
        }
      }
      if (!REDUCE) {
        const int nb1 = p.kkNb1[kk];
        for (int b2 = 0; b2 < nb2; ++b2)
          p.out[p.rowOutOff[ri.row] + b1p + size_t(nb1) * (b2 + size_t(nb2) * sub)] = v[b2];
      }
    }
    if (REDUCE) {
      // symmetrizeCoupling along b2 (scattering.cpp:1250-1284), in registers
      if (modeOn) {
        const double* e2r = se2 + r * p.nb2max;
        for (int b = 0; b < nb2;) {
          int deg = 0;
          for (int i = b; i < nb2; ++i) {
            if (fabs(e2r[b] - e2r[i]) < 1.0e-6) deg++;
            else break;
          }
          if (deg > 1) {
            double s = 0.;
#pragma unroll
            for (int i = 0; i < kSepNb2; ++i)
              if (i >= b && i < b + deg) s += v[i] / double(float(deg));
#pragma unroll
            for (int i = 0; i < kSepNb2; ++i)
              if (i >= b && i < b + deg) v[i] = s;
          }
          b += deg;
        }
      }
      // ... and along nu (scattering.cpp:1286-1316), through LDS when the modes are degenerate
      int mydeg = 1, mystart = sub;
      if (modeOn) {
        const double* wr = sw3 + r * nm;
        for (int s0 = 0; s0 < nm;) {
          int deg = 0;
          for (int i = s0; i < nm; ++i) {
            if (fabs(wr[s0] - wr[i]) < 1.0e-6) deg++;
            else break;
          }
          if (sub >= s0 && sub < s0 + deg) {
            mystart = s0;
            mydeg = deg;
          }
          s0 += deg;
        }
      }
      if (__syncthreads_or(mydeg > 1)) {
        if (modeOn)
          for (int b2 = 0; b2 < nb2; ++b2) gs[(r * nm + sub) * p.nb2max + b2] = v[b2];
        __syncthreads();
        if (modeOn && mydeg > 1)
          for (int b2 = 0; b2 < nb2; ++b2) {
            double s = 0.;
            for (int i = 0; i < mydeg; ++i) s += gs[(r * nm + mystart + i) * p.nb2max + b2] / double(float(mydeg));
#pragma unroll
            for (int b = 0; b < kSepNb2; ++b)
              if (b == b2) v[b] = s;
          }
      }
      // Gaussian-delta weights (el_scattering.cpp:277-380)
      if (modeOn && w3 >= p.cutoff) {
#pragma unroll
        for (int b2 = 0; b2 < kSepNb2; ++b2) {
          if (b2 >= nb2) break;
          const double en2 = se2[r * p.nb2max + b2];
          const double cpl = v[b2] * inv_g;
          const int is2 = ri.soff + b2;
#pragma unroll
          for (int gi = 0; gi < kGMax; ++gi) {
            if (gi >= gsize) break;
            double d1, d2;
            if (p.adaptive) {
              const AdaptiveWidth aw = adaptive_width(p, int(e1 - p.e1pool) + gi, is2);
              d1 = adaptive_delta(e1[gi] - en2 + w3, aw);
              d2 = adaptive_delta(e1[gi] - en2 - w3, aw);
            } else {
              d1 = gaussian_delta(e1[gi] - en2 + w3, p.invSigma, p.deltaPref);
              d2 = gaussian_delta(e1[gi] - en2 - w3, p.invSigma, p.deltaPref);
            }
            if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
              if (c >= nCalc) break;
              const double f2 = p.fermi[size_t(c) * p.numStates + is2];
              part[gi][c] += cpl * ((f2 + n3[c]) * d1 + (1. - f2 + n3[c]) * d2) * p.norm / w3 *
                             3.14159265358979323846;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (REDUCE) reduce_partials<CM>(p, L, smem, part);
}

inline size_t sep_x_lds_bytes(int elem, int nG0, int ncol, int m0, int nw, int nm, int nb2max) {
  const size_t t2 = 2 * size_t(elem);
  return (size_t(nG0) * ncol * t2 + size_t(m0) * nG0 * t2 + 15) / 16 * 16 +
         size_t(kSepRows) * (ncol + nw * nb2max + nm * nb2max) * sizeof(double2) +
         size_t(kSepRows) * (nm * nb2max + nb2max + nm) * sizeof(double) +
         size_t(m0) * sizeof(SepXShared) + size_t(m0 + 1) * sizeof(int) + 64;
}

// ---- last stage on MFMA
// Same result as k_sep_x.  Per workgroup (one (jy, jz) line of mesh points and
// one panel), the kept rows jx are compacted, then for every tile of 16 of them
//   g3(jx, c) = sum_g w_x^{jx n0(g)} T2(g, jy, jz, c)
// runs as two real GEMMs on v_mfma_*_16x16x4 in the transform precision,
// sharing B(2g + s, c) = (Re, Im) X(g, c):
//   Re g3 = A1 B with A1(m, 2g + s) = (Re w, -Im w),  Im g3 = A2 B with A2 = (Im w, Re w),
// w = w(jx_m, g) from a twiddle table in LDS, B straight from T2 (L2).  The epilogue gives
// TPR threads to each row (thread sub = mode nu).  Workgroups are ordered
// with the panels innermost, so that consecutive workgroups share the rows'
// U2 and the q-line's phonons in L2.
// kXmNt: complex column tiles (16 columns) per wave: 3 for nm * nw <= 192, 6 for <= 384
constexpr int kXmKsc = 3;   // MFMA k-steps per B chunk

struct SepXmRow {            // per kept row, staged in LDS
  int jx, row, iq, nb2, u2off, soff;
};

typedef short s16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// Orders LDS accesses between the lanes of one wavefront (no block barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row data of a 16-row tile into LDS: U2 (row r at sU2 + r * u2s, nw x nb2max),
// e2 (16 x nb2max), omega (16 x nm) and the electron occupations of the rows'
// states (nCalc x 16 x nb2max).  Every global load of a chunk is issued before
// the LDS stores, so a tile waits for one memory round trip, not one per
// element.  Needs 16 nb2max <= 256, 16 nm <= 512 and nCalc 16 nb2max <= 512.
__device__ inline void sep_xm_stage_rows(const CoupParams& p, const SepXmRow* rows, int m0r, int nr,
                                         int nw, int nm, int u2s, double2* sU2, double* se2,
                                         double* sw3, int* sg3, double* sf2) {
  const int t = threadIdx.x, nb2max = p.nb2max;
  const int ne = 16 * nb2max, nwv = 16 * nm, nf = p.nCalc * ne;
  double ev = 0., wv[2] = {0., 0.}, fv[2] = {0., 0.};
  int gv[2] = {0, 0};
  if (t < ne) {
    const int r = t / nb2max, b = t % nb2max;
    if (r < nr && b < rows[m0r + r].nb2) ev = p.e2pool[rows[m0r + r].soff + b];
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = t + 256 * u;
    if (i < nwv) {
      const int r = i / nm, nu = i % nm;
      if (r < nr) {
        wv[u] = p.Wpool[size_t(rows[m0r + r].iq) * nm + nu];
        gv[u] = p.qModeGrp[size_t(rows[m0r + r].iq) * nm + nu];
      }
    }
    if (i < nf) {
      const int c = i / ne, o = i % ne, r = o / nb2max, b = o % nb2max;
      if (r < nr && b < rows[m0r + r].nb2)
        fv[u] = p.fermi[size_t(c) * p.numStates + rows[m0r + r].soff + b];
    }
  }
  const int per = nw * nb2max, nU = 16 * per;
  for (int i0 = t; i0 < nU; i0 += 4 * 256) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 256 * u;
      v[u] = make_double2(0., 0.);
      if (i < nU) {
        const int r = i / per, o = i % per;
        if (r < nr) {
          const SepXmRow& rr = rows[m0r + r];
          if (o < nw * rr.nb2) v[u] = p.U2pool[rr.u2off + o];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 256 * u;
      if (i < nU) sU2[(i / per) * u2s + i % per] = v[u];
    }
  }
  if (t < ne) se2[t] = ev;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = t + 256 * u;
    if (i < nwv) {
      sw3[i] = wv[u];
      sg3[i] = gv[u];
    }
    if (i < nf) sf2[i] = fv[u];
  }
}

// Split-BF16 operands of v_mfma_f32_16x16x32_bf16: x ~ hi + lo with
// hi = bf16(x), lo = bf16(x - hi).
__device__ inline void bf16_split8(const float (&x)[8], bf16x8_t& hi, bf16x8_t& lo) {
  s16x8_t h8, l8;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const short h = bf16_bits(x[j]);
    h8[j] = h;
    l8[j] = bf16_bits(x[j] - bf16_value(h));
  }
  hi = __builtin_bit_cast(bf16x8_t, h8);
  lo = __builtin_bit_cast(bf16x8_t, l8);
}

// Second half of an epilogue pass of k_sep_ep (k_sep_xm keeps the same steps
// inline, where the compiler allocates them without spills).  Once the
// wave's g'(nu, b2) rows are in gp (row rl at gp + rl * gps, (nu, b2) at
// nu * nb2max + b2): the phonon rotation, |g|^2, symmetrizeCoupling along b2
// and nu, and the Gaussian-delta weights into part (or the store of |g|^2).
// The thread serves row r (rl within the pass) and mode sub; the row data
// (e2, mode groups, occupations) of the 16-row tile and the b1 group's
// energies se1 are in LDS.  Wave-level ordering only.
template <typename T, bool REDUCE, int NMX, int CM>
__device__ __forceinline__ void sep_epilogue_tail(const CoupParams& p, const SepXmRow& ri, int r, int rl, int sub,
                                                  int nb2, bool modeOn, double w3, double2* gp, int gps,
                                                  const double* se2, const int* sg3, const double* sf2,
                                                  const double* se1, const typename Vec2<T>::type (&e3c)[NMX],
                                                  const double (&n3)[CM], double (&part)[kGMax][CM],
                                                  const double* e1, int kk, int b1p, int gsize, double inv_g) {
  const int nm = p.nm, nCalc = p.nCalc;
  const double scrW = p.screen > 0. ? screen_width(p, int(e1 - p.e1pool)) : 0.;
  // g(nu2, b2) = sum_nu e3(nu, nu2) g'(nu, b2), thread nu2 = sub; |g|^2 in registers
  double v[kSepNb2];
#pragma unroll
  for (int b = 0; b < kSepNb2; ++b) v[b] = 0.;
  if (modeOn) {
#pragma unroll
    for (int b2 = 0; b2 < kSepNb2; ++b2) {
      if (b2 < nb2) {
        double sx = 0., sy = 0.;
#pragma unroll
        for (int nu = 0; nu < NMX; ++nu)
          if (nu < nm) {
            const double2 x = gp[rl * gps + nu * p.nb2max + b2];
            const double er = double(e3c[nu].x), ei = double(e3c[nu].y);
            sx += x.x * er - x.y * ei;
            sy += x.x * ei + x.y * er;
          }
        v[b2] = sx * sx + sy * sy;
      }
    }
    if (!REDUCE) {
      const int nb1 = p.kkNb1[kk];
      for (int b2 = 0; b2 < nb2; ++b2)
        p.out[p.rowOutOff[ri.row] + b1p + size_t(nb1) * (b2 + size_t(nb2) * sub)] = v[b2];
    }
  }
  if (REDUCE) {
    // symmetrizeCoupling along b2 (scattering.cpp:1250-1284), in registers
    if (modeOn) {
      const double* e2r = se2 + r * p.nb2max;
      for (int b = 0; b < nb2;) {
        int deg = 0;
        for (int i = b; i < nb2; ++i) {
          if (fabs(e2r[b] - e2r[i]) < 1.0e-6) deg++;
          else break;
        }
        if (deg > 1) {
          double s2 = 0.;
#pragma unroll
          for (int i = 0; i < kSepNb2; ++i)
            if (i >= b && i < b + deg) s2 += v[i] / double(float(deg));
#pragma unroll
          for (int i = 0; i < kSepNb2; ++i)
            if (i >= b && i < b + deg) v[i] = s2;
        }
        b += deg;
      }
    }
    // ... and along nu (scattering.cpp:1286-1316), through LDS when modes are degenerate
    // the mode's degenerate group (consecutive modes within 1e-6 of the
    // group's first), tabulated per q at set-up
    int mydeg = 1, mystart = sub;
    if (modeOn) {
      const int grp = sg3[r * nm + sub];
      mystart = grp & 0xff;
      mydeg = grp >> 8;
    }
    if (__any(mydeg > 1)) {  // g' of the row is dead here: its slots hold |g|^2
      double* gs = reinterpret_cast<double*>(gp);
      wave_sync();
      if (modeOn)
        for (int b2 = 0; b2 < nb2; ++b2) gs[2 * (rl * gps + sub * p.nb2max + b2)] = v[b2];
      wave_sync();
      if (modeOn && mydeg > 1)
        for (int b2 = 0; b2 < nb2; ++b2) {
          double s2 = 0.;
          for (int i = 0; i < mydeg; ++i)
            s2 += gs[2 * (rl * gps + (mystart + i) * p.nb2max + b2)] / double(float(mydeg));
#pragma unroll
          for (int b = 0; b < kSepNb2; ++b)
            if (b == b2) v[b] = s2;
        }
    }
    // Gaussian-delta weights (el_scattering.cpp:277-380); b2 loop kept rolled
    if (modeOn && w3 >= p.cutoff) {
#pragma unroll 1
      for (int b2 = 0; b2 < nb2; ++b2) {
        double vb = 0.;
#pragma unroll
        for (int b = 0; b < kSepNb2; ++b)
          if (b == b2) vb = v[b];
        const double en2 = se2[r * p.nb2max + b2];
        const double cpl = vb * inv_g;
        const int is2 = ri.soff + b2;
#pragma unroll
        for (int gi = 0; gi < kGMax; ++gi) {
          if (gi >= gsize) break;
          double d1, d2;
          if (p.adaptive) {
            const AdaptiveWidth aw = adaptive_width(p, int(e1 - p.e1pool) + gi, is2);
            d1 = adaptive_delta(se1[gi] - en2 + w3, aw);
            d2 = adaptive_delta(se1[gi] - en2 - w3, aw);
          } else {
            const double x1 = se1[gi] - en2 + w3, x2 = se1[gi] - en2 - w3;
            // with the energy screen on, a pair whose both weights are below
            // exp(-screen^2) of the peak is skipped, as whole rows are
            if (p.screen > 0. && fmin(fabs(x1), fabs(x2)) * p.invSigma >= scrW)
              continue;
            d1 = gaussian_delta(x1, p.invSigma, p.deltaPref);
            d2 = gaussian_delta(x2, p.invSigma, p.deltaPref);
          }
          if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
          for (int c = 0; c < CM; ++c) {
            if (c >= nCalc) break;
            const double f2 = sf2[(c * 16 + r) * p.nb2max + b2];
            part[gi][c] += cpl * ((f2 + n3[c]) * d1 + (1. - f2 + n3[c]) * d2) * p.norm / w3 *
                           3.14159265358979323846;
          }
        }
      }
    }
  }
}

// Second half of an epilogue pass of k_sep_ep, with |g|^2 kept in LDS
// rather than in registers: gsq holds the rpp rows' |g(nu, b2)|^2, (row, nu,
// b2) at (r nm + nu) nb2max + b2; occupations at sf2[(c rpp + r) nb2max + b].  The phonon rotation, symmetrizeCoupling
// along b2 (each thread its own entries) and then along nu (read on the fly
// from the group's entries), and the Gaussian-delta weights into part.  The
// sums run in the same order as sep_epilogue_tail's.  Wave-level ordering.
template <typename T, int NMX, int CM, int GM>
__device__ __forceinline__ void sep_epilogue_tail_lds(const CoupParams& p, const SepXmRow& ri, int r, int rl,
                                                      int sub, int nb2, bool modeOn, double w3, const double2* gp,
                                                      int gps, double* gsq, const double* se2, const int* sg3,
                                                      const double* sf2, const double* se1,
                                                      const typename Vec2<T>::type (&e3c)[NMX],
                                                      const double (&n3)[CM], double (&part)[GM][CM],
                                                      const double* e1, int gsize, double inv_g, int rpp) {
  const int nm = p.nm, nCalc = p.nCalc, nb2max = p.nb2max;
  const double scrW = p.screen > 0. ? screen_width(p, int(e1 - p.e1pool)) : 0.;
  double* own = gsq + (size_t(r) * nm + sub) * nb2max;
  // g(nu2, b2) = sum_nu e3(nu, nu2) g'(nu, b2), thread nu2 = sub
  if (modeOn)
    for (int b2 = 0; b2 < nb2; ++b2) {
      double sx = 0., sy = 0.;
#pragma unroll
      for (int nu = 0; nu < NMX; ++nu)
        if (nu < nm) {
          const double2 x = gp[rl * gps + nu * nb2max + b2];
          const double er = double(e3c[nu].x), ei = double(e3c[nu].y);
          sx += x.x * er - x.y * ei;
          sy += x.x * ei + x.y * er;
        }
      own[b2] = sx * sx + sy * sy;
    }
  // symmetrizeCoupling along b2 (scattering.cpp:1250-1284), own entries
  if (modeOn) {
    const double* e2r = se2 + r * nb2max;
    for (int b = 0; b < nb2;) {
      int deg = 0;
      for (int i = b; i < nb2; ++i) {
        if (fabs(e2r[b] - e2r[i]) < 1.0e-6) deg++;
        else break;
      }
      if (deg > 1) {
        double s2 = 0.;
        for (int i = 0; i < deg; ++i) s2 += own[b + i] / double(float(deg));
        for (int i = 0; i < deg; ++i) own[b + i] = s2;
      }
      b += deg;
    }
  }
  wave_sync();
  // ... along nu (scattering.cpp:1286-1316) and the Gaussian-delta weights
  // (el_scattering.cpp:277-380)
  if (!modeOn || w3 < p.cutoff) return;
  const int grp = sg3[r * nm + sub];
  const int mystart = grp & 0xff, mydeg = grp >> 8;
  const double* g0 = gsq + (size_t(r) * nm + mystart) * nb2max;
#pragma unroll 1
  for (int b2 = 0; b2 < nb2; ++b2) {
    double vb;
    if (mydeg > 1) {
      vb = 0.;
      for (int i = 0; i < mydeg; ++i) vb += g0[size_t(i) * nb2max + b2] / double(float(mydeg));
    } else {
      vb = own[b2];
    }
    const double en2 = se2[r * nb2max + b2];
    const double cpl = vb * inv_g;
    const int is2 = ri.soff + b2;
#pragma unroll
    for (int gi = 0; gi < GM; ++gi) {
      if (gi >= gsize) break;
      double d1, d2;
      if (p.adaptive) {
        const AdaptiveWidth aw = adaptive_width(p, int(e1 - p.e1pool) + gi, is2);
        d1 = adaptive_delta(se1[gi] - en2 + w3, aw);
        d2 = adaptive_delta(se1[gi] - en2 - w3, aw);
      } else {
        const double x1 = se1[gi] - en2 + w3, x2 = se1[gi] - en2 - w3;
        if (p.screen > 0. && fmin(fabs(x1), fabs(x2)) * p.invSigma >= scrW) continue;
        d1 = gaussian_delta(x1, p.invSigma, p.deltaPref);
        d2 = gaussian_delta(x2, p.invSigma, p.deltaPref);
      }
      if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= nCalc) break;
        const double f2 = sf2[(c * rpp + r) * nb2max + b2];
        part[gi][c] += cpl * ((f2 + n3[c]) * d1 + (1. - f2 + n3[c]) * d2) * p.norm / w3 * 3.14159265358979323846;
      }
    }
  }
}

// BF3: the x-stage GEMMs in split-BF16x3 (hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x32_bf16, FP32 accumulation); T must be float.
// BREG: the B operands of all k-steps (at most 12, i.e. 24 x-groups) are loaded
// once per work item, in the prologue, where their latency overlaps the row
// compaction; otherwise they are loaded per tile in chunks of KSB k-steps.
template <typename T, bool REDUCE, int NMX, int CM, bool BF3 = false, int kXmNt = 3, bool BREG = false>
__global__ void __launch_bounds__(256, 2) k_sep_xm(CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  typedef typename Mfma4<T>::acc_t acc_t;
  constexpr int TPR = NMX <= 16 ? 16 : 32;  // epilogue threads per row
  constexpr int RPP = 256 / TPR;            // rows per epilogue pass
  constexpr int KSB = (sizeof(T) == 4 && kXmNt <= 3) ? 12 : 3;  // k-steps of B held in registers
  const int nLinesYZ = p.m1 * p.m2;
  const int nTilesTot = p.nPanels * nLinesYZ;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int Lx = (bid % 8) * perXcd + bid / 8;
  if (Lx >= nTilesTot) return;
  const int yz = Lx / p.nPanels, panel = Lx % p.nPanels;
  const int slot = panel * nLinesYZ + yz;  // partial-sum slot (panel-major, as k_collect reads it)
  const int jy = yz % p.m1, jz = yz / p.m1;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15, lg = lane >> 4;
  if (p.gate && *p.gate == 0) return;
  if (p.slotOn && !p.slotOn[slot]) {
    if (REDUCE && p.slotZero && t < kGMax * kCMax) p.partial[size_t(slot) * kGMax * kCMax + t] = 0.;
    return;
  }
  // no row of the line passes the energy screen: zero partials, before any slab load
  if (p.keepMask && p.keepMask[slot] == 0ull) {
    if (REDUCE && t < kGMax * kCMax) p.partial[size_t(slot) * kGMax * kCMax + t] = 0.;
    return;
  }
  const int nm = p.nm, nw = p.nw, ncol = nm * nw;
  const int NWP = 2 * nw + 1, LDG = nm * NWP;  // g3 row layout: (nu, 2 w2 + ri), padded per nu
  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];
  const int nG0 = p.nG0, nG0p = nG0 + 1;
  const int kkMesh = p.kkMesh[kk];

  // LDS carve-up
  SepXmRow* rows = reinterpret_cast<SepXmRow*>(smem);                          // m0
  int* wcnt = reinterpret_cast<int*>(rows + p.m0);                             // 8
  int* sg0 = wcnt + 8;                                                         // nG0
  T2v* stw = reinterpret_cast<T2v*>(smem + (size_t(p.m0) * sizeof(SepXmRow) + (8 + nG0) * sizeof(int) + 15) / 16 * 16);
  T2v* Wk = stw + p.m0;                                                        // m0 x nG0p
  unsigned char* dyn = reinterpret_cast<unsigned char*>(Wk + size_t(p.m0) * nG0p);
  dyn = smem + (size_t(dyn - smem) + 15) / 16 * 16;
  // Row strides of U2 and g' padded by one element: the 4 rows a wave serves
  // then start in different LDS banks.
  const int u2s = nw * p.nb2max + 1, gps = nm * p.nb2max + 1;
  T* g3 = reinterpret_cast<T*>(dyn);                                           // 16 x LDG
  double2* sU2 = reinterpret_cast<double2*>(dyn + (size_t(16) * LDG * sizeof(T) + 15) / 16 * 16);  // 16 x u2s
  double2* gp = sU2 + size_t(16) * u2s;                                        // RPP x gps
  double* se2 = reinterpret_cast<double*>(gp + size_t(RPP) * gps);             // 16 x nb2max
  double* sw3 = se2 + size_t(16) * p.nb2max;                                   // 16 x nm
  double* sf2 = sw3 + size_t(16) * nm;                                         // nCalc x 16 x nb2max
  double* se1 = sf2 + size_t(p.nCalc) * 16 * p.nb2max;                         // kGMax
  int* sg3 = reinterpret_cast<int*>(se1 + kGMax);                              // 16 x nm

  // twiddles and group coordinates
  for (int i = t; i < p.m0; i += 256) {
    const double2 d = p.tw0[i];
    stw[i] = T2v{T(d.x), T(d.y)};
  }
  for (int i = t; i < nG0; i += 256) sg0[i] = p.g0val[i];
  if (t < p.panelGSize[panel]) se1[t] = p.e1pool[p.kkStateOff[kk] + p.panelGStart[panel] + t];

  const T2v* X = static_cast<const T2v*>(p.T2) + size_t(panel) * (p.t2PanelStride / 2) +
                 (size_t(jy) * p.m2 + jz) * ncol;
  const size_t gStride = size_t(p.m1) * p.m2 * ncol;  // T2v elements between groups
  const int KS = (2 * nG0 + 3) / 4;                    // k-steps of 4 real K
  constexpr int KSR = BREG ? 12 : 1;
  T bAll[KSR][kXmNt];
  if constexpr (BREG) {
#pragma unroll
    for (int q = 0; q < KSR; ++q) {
      const int k = 4 * q + lg, g = k >> 1, kc = k & 1;
      const bool gOk = g < nG0;
#pragma unroll
      for (int j = 0; j < kXmNt; ++j) {
        const int c = (wave * kXmNt + j) * 16 + lr;
        T v = T(0);  // one component: 4-byte (FP32) loads, half the bytes of a complex load
        if (gOk && c < ncol) v = reinterpret_cast<const T*>(X)[2 * (size_t(g) * gStride + c) + kc];
        bAll[q][j] = v;
      }
    }
  }

  // ---- rows of the line, energy screen, compaction (ascending jx)
  int nkeep = 0;
  if (p.keepMask) {
    // kept rows from the screen kernel: one load per kept row
    const unsigned long long mask = p.keepMask[slot];
    nkeep = __popcll(mask);
    for (int jx = t; jx < p.m0; jx += 256)
      if ((mask >> jx) & 1ull) {
        const int mesh = (jz * p.m1 + jy) * p.m0 + jx;
        const int4 mr = p.meshRow[mesh];
        rows[__popcll(mask & ((1ull << jx) - 1ull))] =
            SepXmRow{jx, mr.x, mesh_diff_index(mesh, kkMesh, p.m0, p.m1, p.m2), mr.y, mr.z, mr.w};
      }
    __syncthreads();
  } else
  for (int j0 = 0; j0 < p.m0; j0 += 256) {
    const int jx = j0 + t;
    SepXmRow ri{jx, -1, 0, 0, 0, 0};    if (jx < p.m0) {
      const int mesh = (jz * p.m1 + jy) * p.m0 + jx;
      ri.iq = mesh_diff_index(mesh, kkMesh, p.m0, p.m1, p.m2);
      double wlo = 0., whi = 0.;
      if (p.screen > 0.) {
        wlo = p.qWlo[ri.iq];
        whi = p.qWhi[ri.iq];
      }
      int row = p.meshToRow[mesh];
      if (row < p.rowLo || row >= p.rowHi) row = -1;
      if (row >= 0) {
        ri.nb2 = p.rowNb2[row];
        ri.u2off = p.rowU2off[row];
        ri.soff = p.rowStateOff[row];
        bool keep = true;
        if (p.screen > 0.) {
          const double2 er = p.rowE2range[row];
          const double e2lo = er.x, e2hi = er.y;
          const double* e1 = p.e1pool + p.kkStateOff[kk] + p.panelGStart[panel];
          const double lim = screen_width(p, int(e1 - p.e1pool)) / p.invSigma;
          keep = false;
          if (whi >= wlo) {  // at least one mode above the cutoff
            for (int gi = 0; gi < p.panelGSize[panel]; ++gi) {
              const double a_lo = e1[gi] - e2hi + wlo, a_hi = e1[gi] - e2lo + whi;
              const double b_lo = e1[gi] - e2hi - whi, b_hi = e1[gi] - e2lo - wlo;
              keep |= (a_hi > -lim && a_lo < lim) || (b_hi > -lim && b_lo < lim);
            }
          }
        }
        if (keep) ri.row = row;
      }
    }
    const bool keep = ri.row >= 0;
    const unsigned long long mask = __ballot(keep);
    if (lane == 0) wcnt[wave] = __popcll(mask);
    __syncthreads();
    int off = nkeep, tot = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wave) off += wcnt[w];
      tot += wcnt[w];
    }
    if (keep) rows[off + __popcll(mask & ((1ull << lane) - 1ull))] = ri;
    nkeep += tot;
    __syncthreads();
  }
  if (t == 0 && p.keptRows && nkeep > 0) atomicAdd(p.keptRows, (unsigned long long)nkeep);
  if (nkeep == 0) {
    if (REDUCE) {
      double part[kGMax][CM];
      for (int gi = 0; gi < kGMax; ++gi)
        for (int c = 0; c < CM; ++c) part[gi][c] = 0.;
      reduce_partials<CM>(p, slot, smem, part);
    }
    return;
  }
  // twiddles of the kept rows: Wk(k, g) = w_x^{jx_k n0(g)}, zero column g = nG0 pads the K loop
  for (int i = t; i < nkeep * nG0p; i += 256) {
    const int k = i / nG0p, g = i % nG0p;
    Wk[i] = g < nG0 ? stw[(rows[k].jx * sg0[g]) % p.m0] : T2v{T(0), T(0)};
  }
  __syncthreads();

  double part[kGMax][CM];
#pragma unroll
  for (int gi = 0; gi < kGMax; ++gi)
#pragma unroll
    for (int c = 0; c < CM; ++c) part[gi][c] = 0.;
  const int nCalc = p.nCalc;
  const int gstart = p.panelGStart[panel], gsize = p.panelGSize[panel];
  const double* e1 = p.e1pool + p.kkStateOff[kk] + gstart;
  const double scrW = p.screen > 0. ? screen_width(p, int(e1 - p.e1pool)) : 0.;
  const double inv_g = 1. / double(float(gsize));

  for (int m0r = 0; m0r < nkeep; m0r += 16) {
    const int nr = min(16, nkeep - m0r);
    const bool rowOk = lr < nr;
    const T2v* wrow = Wk + size_t(m0r + (rowOk ? lr : 0)) * nG0p;
    // phonon eigenvector column, Bose factors of the epilogue row (r, mode sub)
    T2v e3c[NMX];
    double n3[CM];
    auto load_e3 = [&](int pr) {
      const int r = pr + t / TPR, sub = t % TPR;
#pragma unroll
      for (int c = 0; c < CM; ++c) n3[c] = 0.;
      if (r < nr && sub < nm) {
        const int iq = rows[m0r + r].iq;
        const T2v* e3 = static_cast<const T2v*>(p.E3T) + size_t(iq) * nm * nm + sub;
#pragma unroll
        for (int nu = 0; nu < NMX; ++nu)
          if (nu < nm) e3c[nu] = e3[nu * nm];
#pragma unroll
        for (int c = 0; c < CM; ++c)
          if (c < p.nCalc) n3[c] = p.bose[(size_t(iq) * nm + sub) * p.nCalc + c];
      }
    };
    // Cr = A1 B, Ci = A2 B with B(2g + s, c) = (Xr, Xi)(g, c), A1(m, 2g + s) = (Wr, -Wi),
    // A2(m, 2g + s) = (Wi, Wr); wave owns complex column tiles wave*kXmNt .. +kXmNt
    acc_t accr[kXmNt], acci[kXmNt];
#pragma unroll
    for (int j = 0; j < kXmNt; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) accr[j][i] = acci[j][i] = T(0);
    if constexpr (BF3) {
      sep_xm_stage_rows(p, rows, m0r, nr, nw, nm, u2s, sU2, se2, sw3, sg3, sf2);
      // gfx950's v_mfma_f32_16x16x32_bf16, 32 real K per step: lane (lr, lg)
      // holds groups g = 16 kb + 4 lg + h, h < 4, (Re, Im) each, in the A and
      // the B operand alike (the pairing of K slots is all the product needs)
      for (int kb = 0; kb < (2 * nG0 + 31) / 32; ++kb) {
        const int gA = 16 * kb + 4 * lg;
        float a1[8], a2[8];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const T2v w = wrow[gA + h < nG0 ? gA + h : nG0];
          const bool on = rowOk && gA + h < nG0;
          a1[2 * h] = on ? w.x : 0.f;
          a1[2 * h + 1] = on ? -w.y : 0.f;
          a2[2 * h] = on ? w.y : 0.f;
          a2[2 * h + 1] = on ? w.x : 0.f;
        }
        bf16x8_t a1h, a1l, a2h, a2l;
        bf16_split8(a1, a1h, a1l);
        bf16_split8(a2, a2h, a2l);
#pragma unroll
        for (int j = 0; j < kXmNt; ++j) {
          const int c = (wave * kXmNt + j) * 16 + lr;
          float b[8];
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            T2v x = T2v{0.f, 0.f};
            if (c < ncol && gA + h < nG0) x = X[size_t(gA + h) * gStride + c];
            b[2 * h] = x.x;
            b[2 * h + 1] = x.y;
          }
          bf16x8_t bh, bl;
          bf16_split8(b, bh, bl);
          accr[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1h, bh, accr[j], 0, 0, 0);
          accr[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1h, bl, accr[j], 0, 0, 0);
          accr[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1l, bh, accr[j], 0, 0, 0);
          acci[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h, bh, acci[j], 0, 0, 0);
          acci[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h, bl, acci[j], 0, 0, 0);
          acci[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2l, bh, acci[j], 0, 0, 0);
        }
      }
    } else if constexpr (BREG) {
      sep_xm_stage_rows(p, rows, m0r, nr, nw, nm, u2s, sU2, se2, sw3, sg3, sf2);
      // fully unrolled with a predicate (no early exit), so that bAll stays in registers
#pragma unroll
      for (int q = 0; q < KSR; ++q) {
        if (q < KS) {
          const int k = 4 * q + lg, g = k >> 1, kc = k & 1;
          const bool gOk = g < nG0;
          const T2v w = wrow[gOk ? g : nG0];
          const bool on = rowOk && gOk;
          const T a1 = on ? (kc ? -w.y : w.x) : T(0);
          const T a2 = on ? (kc ? w.x : w.y) : T(0);
#pragma unroll
          for (int j = 0; j < kXmNt; ++j) {
            accr[j] = Mfma4<T>::mac(a1, bAll[q][j], accr[j]);
            acci[j] = Mfma4<T>::mac(a2, bAll[q][j], acci[j]);
          }
        }
      }
    } else
    for (int kb = 0; kb < KS; kb += KSB) {
      // B operands of KSB k-steps straight from T2, then the MFMAs
      T bv[KSB][kXmNt];
#pragma unroll
      for (int q = 0; q < KSB; ++q) {
        const int k = 4 * (kb + q) + lg, g = k >> 1, kc = k & 1;
        const bool gOk = g < nG0 && kb + q < KS;
#pragma unroll
        for (int j = 0; j < kXmNt; ++j) {
          const int c = (wave * kXmNt + j) * 16 + lr;
          T v = T(0);  // one component of the complex element
          if (gOk && c < ncol) v = reinterpret_cast<const T*>(X)[2 * (size_t(g) * gStride + c) + kc];
          bv[q][j] = v;
        }
      }
      // the tile's row data into LDS while the B loads are in flight; the
      // previous tile's epilogue ended with a barrier
      if (kb == 0) sep_xm_stage_rows(p, rows, m0r, nr, nw, nm, u2s, sU2, se2, sw3, sg3, sf2);
#pragma unroll
      for (int q = 0; q < KSB; ++q) {
        if (kb + q < KS) {
          const int k = 4 * (kb + q) + lg, g = k >> 1, kc = k & 1;
          const bool gOk = g < nG0;
          const T2v w = wrow[gOk ? g : nG0];
          const bool on = rowOk && gOk;
          const T a1 = on ? (kc ? -w.y : w.x) : T(0);
          const T a2 = on ? (kc ? w.x : w.y) : T(0);
#pragma unroll
          for (int j = 0; j < kXmNt; ++j) {
            accr[j] = Mfma4<T>::mac(a1, bv[q][j], accr[j]);
            acci[j] = Mfma4<T>::mac(a2, bv[q][j], acci[j]);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kXmNt; ++j) {
      const int c = (wave * kXmNt + j) * 16 + lr;
      if (c < ncol) {
        const int nu = c / nw, w2 = c - nu * nw;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          T* dst = g3 + Mfma4<T>::row(lg, i) * LDG + nu * NWP + 2 * w2;
          dst[0] = accr[j][i];
          dst[1] = acci[j][i];
        }
      }
    }
    // the first pass's phonon data: loads in flight across the barrier and the
    // electron rotation, registers free during the staging and the MFMAs
    load_e3(0);
    __syncthreads();

    // ---- epilogue, RPP rows per pass (the first pass's phonon data was prefetched)
    for (int pr = 0; pr < 16; pr += RPP) {
      const int r = pr + t / TPR, sub = t % TPR, rl = t / TPR;
      const bool rowOn = r < nr;
      const SepXmRow ri = rows[m0r + (rowOn ? r : 0)];
      const int nb2 = rowOn ? ri.nb2 : 0;
      const bool modeOn = rowOn && sub < nm;
      if (pr > 0) load_e3(pr);
      const double w3 = modeOn ? sw3[r * nm + sub] : 0.;
      // g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2), thread nu = sub.  Each wave
      // owns whole rows from here on: only wave-level ordering is needed.
      if (modeOn) {
        const T* gr = g3 + r * LDG + sub * NWP;
        for (int b2 = 0; b2 < nb2; ++b2) {
          const double2* u = sU2 + size_t(r) * u2s + b2 * nw;
          double sx = 0., sy = 0.;
          for (int w2 = 0; w2 < nw; ++w2) {
            const double xr = double(gr[2 * w2]), xi = double(gr[2 * w2 + 1]);
            sx += u[w2].x * xr + u[w2].y * xi;
            sy += u[w2].x * xi - u[w2].y * xr;
          }
          gp[rl * gps + sub * p.nb2max + b2] = make_double2(sx, sy);
        }
      }
      wave_sync();
      // g(nu2, b2) = sum_nu e3(nu, nu2) g'(nu, b2), thread nu2 = sub; |g|^2 in registers
      double v[kSepNb2];
#pragma unroll
      for (int b = 0; b < kSepNb2; ++b) v[b] = 0.;
      if (modeOn) {
#pragma unroll
        for (int b2 = 0; b2 < kSepNb2; ++b2) {
          if (b2 < nb2) {
            double sx = 0., sy = 0.;
#pragma unroll
            for (int nu = 0; nu < NMX; ++nu)
              if (nu < nm) {
                const double2 x = gp[rl * gps + nu * p.nb2max + b2];
                const double er = double(e3c[nu].x), ei = double(e3c[nu].y);
                sx += x.x * er - x.y * ei;
                sy += x.x * ei + x.y * er;
              }
            v[b2] = sx * sx + sy * sy;
          }
        }
        if (!REDUCE) {
          const int nb1 = p.kkNb1[kk];
          for (int b2 = 0; b2 < nb2; ++b2)
            p.out[p.rowOutOff[ri.row] + b1p + size_t(nb1) * (b2 + size_t(nb2) * sub)] = v[b2];
        }
      }
      if (REDUCE) {
        // symmetrizeCoupling along b2 (scattering.cpp:1250-1284), in registers
        if (modeOn) {
          const double* e2r = se2 + r * p.nb2max;
          for (int b = 0; b < nb2;) {
            int deg = 0;
            for (int i = b; i < nb2; ++i) {
              if (fabs(e2r[b] - e2r[i]) < 1.0e-6) deg++;
              else break;
            }
            if (deg > 1) {
              double s2 = 0.;
#pragma unroll
              for (int i = 0; i < kSepNb2; ++i)
                if (i >= b && i < b + deg) s2 += v[i] / double(float(deg));
#pragma unroll
              for (int i = 0; i < kSepNb2; ++i)
                if (i >= b && i < b + deg) v[i] = s2;
            }
            b += deg;
          }
        }
        // ... and along nu (scattering.cpp:1286-1316), through LDS when modes are degenerate
        // the mode's degenerate group (consecutive modes within 1e-6 of the
        // group's first), tabulated per q at set-up
        int mydeg = 1, mystart = sub;
        if (modeOn) {
          const int grp = sg3[r * nm + sub];
          mystart = grp & 0xff;
          mydeg = grp >> 8;
        }
        if (__any(mydeg > 1)) {  // g' of the row is dead here: its slots hold |g|^2
          double* gs = reinterpret_cast<double*>(gp);
          wave_sync();
          if (modeOn)
            for (int b2 = 0; b2 < nb2; ++b2) gs[2 * (rl * gps + sub * p.nb2max + b2)] = v[b2];
          wave_sync();
          if (modeOn && mydeg > 1)
            for (int b2 = 0; b2 < nb2; ++b2) {
              double s2 = 0.;
              for (int i = 0; i < mydeg; ++i)
                s2 += gs[2 * (rl * gps + (mystart + i) * p.nb2max + b2)] / double(float(mydeg));
#pragma unroll
              for (int b = 0; b < kSepNb2; ++b)
                if (b == b2) v[b] = s2;
            }
        }
        // Gaussian-delta weights (el_scattering.cpp:277-380); b2 loop kept rolled
        if (modeOn && w3 >= p.cutoff) {
#pragma unroll 1
          for (int b2 = 0; b2 < nb2; ++b2) {
            double vb = 0.;
#pragma unroll
            for (int b = 0; b < kSepNb2; ++b)
              if (b == b2) vb = v[b];
            const double en2 = se2[r * p.nb2max + b2];
            const double cpl = vb * inv_g;
            const int is2 = ri.soff + b2;
#pragma unroll
            for (int gi = 0; gi < kGMax; ++gi) {
              if (gi >= gsize) break;
              double d1, d2;
              if (p.adaptive) {
                const AdaptiveWidth aw = adaptive_width(p, int(e1 - p.e1pool) + gi, is2);
                d1 = adaptive_delta(se1[gi] - en2 + w3, aw);
                d2 = adaptive_delta(se1[gi] - en2 - w3, aw);
              } else {
                const double x1 = se1[gi] - en2 + w3, x2 = se1[gi] - en2 - w3;
                // with the energy screen on, a pair whose both weights are below
                // exp(-screen^2) of the peak is skipped, as whole rows are
                if (p.screen > 0. && fmin(fabs(x1), fabs(x2)) * p.invSigma >= scrW)
                  continue;
                d1 = gaussian_delta(x1, p.invSigma, p.deltaPref);
                d2 = gaussian_delta(x2, p.invSigma, p.deltaPref);
              }
              if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
              for (int c = 0; c < CM; ++c) {
                if (c >= nCalc) break;
                const double f2 = sf2[(c * 16 + r) * p.nb2max + b2];
                part[gi][c] += cpl * ((f2 + n3[c]) * d1 + (1. - f2 + n3[c]) * d2) * p.norm / w3 *
                               3.14159265358979323846;
              }
            }
          }
        }
      }
      wave_sync();  // the next pass reuses this wave's g' rows
    }
    __syncthreads();  // the next tile rewrites g3, U2, e2, omega
  }
  if (REDUCE) reduce_partials<CM>(p, slot, smem, part);
}


// ------------------------------------------------------------------ split last stage
// The last stage as two kernels with small footprints, for FP32 data, nm, nw
// <= 16, nm nw <= 192 and precomputed row masks: many more workgroups share a
// CU than with the fused k_sep_xm (about 65 KB of LDS and 255 VGPRs), which
// is bound by memory latency.  g3 rows go through HBM, (slot, jx, col), and
// the epilogue is the same sequence of operations on the same values, so the
// results equal the fused kernel's.

// x-stage: g3(row, (nu, w2)) = sum_g w_x^{jx n0(g)} T2(g; nu, w2) of every
// row kept by the slot's mask, stored at G3 (slot, jx) as complex T.
// ROT: the kernel also applies the electron rotation of the epilogue,
// g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2) in FP64 with w2 in order
// (the arithmetic of k_sep_ep2's phase A, on the same FP32 g3 values), and
// stores g' instead: complex FP64 at G3 (slot, jx, b2, nu), nb2max nm entries
// per row of which the row's nb2 nm are written.  At cfg2's 1.7 bands per
// inner point that is about a fifth of the bytes of g3.
template <typename T, int kXmNt, bool ROT = false>
__global__ void __launch_bounds__(256) k_sep_xg(CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  typedef typename Mfma4<T>::acc_t acc_t;
  const int nLinesYZ = p.m1 * p.m2;
  const int nTilesTot = p.nPanels * nLinesYZ;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int Lx = (bid % 8) * perXcd + bid / 8;
  if (Lx >= nTilesTot) return;
  const int yz = Lx / p.nPanels, panel = Lx % p.nPanels;
  const int slot = panel * nLinesYZ + yz;
  const int jy = yz % p.m1, jz = yz / p.m1;
  if (p.gate && *p.gate == 0) return;
  if (p.slotOn && !p.slotOn[slot]) return;
  const unsigned long long mask = p.keepMask[slot];
  if (mask == 0ull) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15, lg = lane >> 4;
  if (t == 0 && p.keptSlots) atomicAdd(p.keptSlots, 1ull);
  const int nm = p.nm, nw = p.nw, ncol = nm * nw;
  const int nG0 = p.nG0, nG0p = nG0 + 1;
  int* jxs = reinterpret_cast<int*>(smem);                                     // 64: kept rows, ascending
  int* snb = jxs + 64;                                                         // ROT: 64 band counts
  int* su2 = snb + 64;                                                         // ROT: 64 U2 offsets
  int* sg0 = su2 + 64;                                                         // nG0
  T2v* stw = reinterpret_cast<T2v*>(smem + (size_t(192 + nG0) * sizeof(int) + 15) / 16 * 16);  // m0
  T2v* Wk = stw + p.m0;                                                        // nkeep x nG0p
  T2v* sG = Wk + size_t(p.m0) * nG0p;  // ROT: one tile's g3, 16 rows x (nm (nw + 1) + 1)
  // ROT: the tile's U2, 16 rows x (nb2max x 16 + 1) complex FP64 (w2 fastest, zero beyond the
  // row's bands and nw), 16-byte aligned after sG
  const int us = p.nb2max * 16 + 1;
  double2* sUx = reinterpret_cast<double2*>(
      smem + (size_t(reinterpret_cast<unsigned char*>(sG + 16 * (nm * (nw + 1) + 1)) - smem) + 15) / 16 * 16);
  const T2v* X = static_cast<const T2v*>(p.T2) + size_t(panel) * (p.t2PanelStride / 2) +
                 (size_t(jy) * p.m2 + jz) * ncol;
  const size_t gStride = size_t(p.m1) * p.m2 * ncol;
  const int KS = (2 * nG0 + 3) / 4;  // <= 12, checked by the launcher
  T bAll[12][kXmNt];
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    const int k = 4 * q + lg, g = k >> 1, kc = k & 1;
    const bool gOk = g < nG0;
#pragma unroll
    for (int j = 0; j < kXmNt; ++j) {
      const int c = (wave * kXmNt + j) * 16 + lr;
      T v = T(0);  // one component: 4-byte (FP32) loads, half the bytes of a complex load
      if (gOk && c < ncol) v = reinterpret_cast<const T*>(X)[2 * (size_t(g) * gStride + c) + kc];
      bAll[q][j] = v;
    }
  }
  for (int i = t; i < p.m0; i += 256) {
    const double2 d = p.tw0[i];
    stw[i] = T2v{T(d.x), T(d.y)};
  }
  for (int i = t; i < nG0; i += 256) sg0[i] = p.g0val[i];
  if (t < p.m0 && ((mask >> t) & 1ull)) {
    const int k = __popcll(mask & ((1ull << t) - 1ull));
    jxs[k] = t;
    if (ROT) {
      const int4 mr = p.meshRow[(jz * p.m1 + jy) * p.m0 + t];
      snb[k] = mr.y;
      su2[k] = mr.z;
    }
  }
  const int nkeep = __popcll(mask);
  __syncthreads();
  for (int i = t; i < nkeep * nG0p; i += 256) {
    const int k = i / nG0p, g = i % nG0p;
    Wk[i] = g < nG0 ? stw[(jxs[k] * sg0[g]) % p.m0] : T2v{T(0), T(0)};
  }
  __syncthreads();
  T2v* out = static_cast<T2v*>(p.G3) + size_t(slot) * p.m0 * ncol;
  // ROT: g' rows of nb2max nm complex FP64, (b2, nu) within a row
  const int gs = nw + 1, rs = nm * gs + 1, rowG = p.nb2max * nm;
  double2* outR = static_cast<double2*>(p.G3) + size_t(slot) * p.m0 * rowG;
  int sgOff[kXmNt];  // ROT: LDS offset of the lane's column (nu, w2) within a tile row
#pragma unroll
  for (int j = 0; j < kXmNt; ++j) {
    const int c = (wave * kXmNt + j) * 16 + lr;
    sgOff[j] = (c / nw) * gs + c % nw;
  }
  for (int m0r = 0; m0r < nkeep; m0r += 16) {
    const int nr = min(16, nkeep - m0r);
    const bool rowOk = lr < nr;
    const T2v* wrow = Wk + size_t(m0r + (rowOk ? lr : 0)) * nG0p;
    acc_t accr[kXmNt], acci[kXmNt];
#pragma unroll
    for (int j = 0; j < kXmNt; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) accr[j][i] = acci[j][i] = T(0);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      if (q < KS) {
        const int k = 4 * q + lg, g = k >> 1, kc = k & 1;
        const bool gOk = g < nG0;
        const T2v w = wrow[gOk ? g : nG0];
        const bool on = rowOk && gOk;
        const T a1 = on ? (kc ? -w.y : w.x) : T(0);
        const T a2 = on ? (kc ? w.x : w.y) : T(0);
#pragma unroll
        for (int j = 0; j < kXmNt; ++j) {
          accr[j] = Mfma4<T>::mac(a1, bAll[q][j], accr[j]);
          acci[j] = Mfma4<T>::mac(a2, bAll[q][j], acci[j]);
        }
      }
    }
    if constexpr (ROT) {
      __syncthreads();  // the previous tile's rotation has read sG
#pragma unroll
      for (int j = 0; j < kXmNt; ++j) {
        const int c = (wave * kXmNt + j) * 16 + lr;
        if (c < ncol) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = Mfma4<T>::row(lg, i);
            if (row < nr) sG[row * rs + sgOff[j]] = T2v{accr[j][i], acci[j][i]};
          }
        }
      }
      // U2 of the tile's rows, each entry read from HBM once (not once per mode)
      for (int i = t; i < 16 * p.nb2max * 16; i += 256) {
        const int r = i / (p.nb2max * 16), o = i - r * (p.nb2max * 16), b2 = o >> 4, w2 = o & 15;
        double2 v = make_double2(0., 0.);
        if (r < nr && b2 < snb[m0r + r] && w2 < nw) v = p.U2pool[su2[m0r + r] + b2 * nw + w2];
        sUx[r * us + o] = v;
      }
      __syncthreads();
      // thread (row r, mode nu), nm <= 16 so 16 rows x nm <= 256 threads
      if (t < nr * nm) {
        const int r = t / nm, nu = t - r * nm, k = m0r + r;
        const int nb2 = snb[k];
        const double2* u = sUx + r * us;
        const T2v* x = sG + r * rs + nu * gs;
        double2* o = outR + size_t(jxs[k]) * rowG + nu;
        for (int b2 = 0; b2 < nb2; ++b2) {
          double ar = 0., ai = 0.;
#pragma unroll
          for (int w2 = 0; w2 < 16; ++w2) {
            if (w2 < nw) {
              const double2 c = u[b2 * 16 + w2];
              const double xr = double(x[w2].x), xi = double(x[w2].y);
              ar += c.x * xr + c.y * xi;
              ai += c.x * xi - c.y * xr;
            }
          }
          o[b2 * nm] = make_double2(ar, ai);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < kXmNt; ++j) {
        const int c = (wave * kXmNt + j) * 16 + lr;
        if (c < ncol) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = Mfma4<T>::row(lg, i);
            if (row < nr) out[size_t(jxs[m0r + row]) * ncol + c] = T2v{accr[j][i], acci[j][i]};
          }
        }
      }
    }
  }
}

// rot: the x-stage with the electron rotation (a tile of g3 and its rows' U2 in LDS)
inline size_t sep_xg_lds_bytes(int elem, int nG0, int m0, int nm = 0, int nw = 0, bool rot = false,
                               int nb2max = 0) {
  size_t b = (size_t(192 + nG0) * sizeof(int) + 15) / 16 * 16 + size_t(m0) * (nG0 + 2) * 2 * elem;
  if (rot) b += (size_t(16) * (nm * (nw + 1) + 1) * 2 * elem + 15) / 16 * 16 +
                size_t(16) * (nb2max * 16 + 1) * sizeof(double2);
  return b + 64;
}

// The y-stage and the x-stage in one kernel, so that the x-stage slabs T2
// never go through HBM.  Workgroup (block of 16 columns, jz, panel): the T1
// entries of every line at (jz, the block's columns) are staged in LDS once;
// then, per chunk of kYxJC y coordinates, the y-stage sums
// T2(g, jy, c) = sum_{lines of group g} w_y^{jy f1} T1(line, jz, c) into LDS
// (FP64, the group's lines in ascending order, as k_sep_y does) and wave w
// runs the x-stage of slots (jy0 + w + 4i, jz) on v_mfma_f32_16x16x4f32 as
// k_sep_xg does (A = the twiddles of the slot's kept rows, B = T2 read from
// LDS), writing the kept rows' g3 for the block's columns (waves take the
// chunk's slots in turn).  The same values
// in the same order as k_sep_y + k_sep_xg, so the same g3; HBM sees T1 once
// and the g3 rows, not T2 twice.
constexpr int kYxJC = 4;  // y coordinates per chunk: one slot per wave (8: 47 KB of LDS, slower)
constexpr int kYxRow = 17;  // LDS row of 16 columns, padded: rows of different groups in different banks

inline size_t sep_yx_lds_bytes(int nLines, int nG0, int m0) {
  size_t b = (size_t(nLines) * kYxRow * sizeof(float2) + 15) / 16 * 16;  // T1 of the block
  b += size_t(nLines) * kYxJC * sizeof(double2);                         // y twiddles of a chunk
  b += size_t(nG0) * kYxJC * kYxRow * sizeof(float2);                    // T2 of a chunk
  b += size_t(m0) * (nG0 + 2) * sizeof(float2);                          // x twiddles of every row
  b += size_t(2 * nLines + nG0 + 1 + 4 * 64) * sizeof(int);          // entries, groups, kept rows
  return b + 64;
}

static __global__ void __launch_bounds__(256) k_sep_yx(SepParams sp, CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Mfma4<float>::acc_t acc_t;
  const int cb = blockIdx.x, jz = blockIdx.y, panel = blockIdx.z;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15, lg = lane >> 4;
  const int ncol = sp.ncol, my = sp.my, mz = sp.mz, m0 = p.m0, nG0 = sp.nG0, nL = sp.nLines;
  const int nG0p = nG0 + 1, wxs = nG0 + 2;  // wxs: padded row of the x twiddles
  float2* sT1 = reinterpret_cast<float2*>(smem);  // (entry in group order, kYxRow)
  double2* sWy = reinterpret_cast<double2*>(smem + (size_t(nL) * kYxRow * sizeof(float2) + 15) / 16 * 16);  // (entry, jj)
  float2* sT2 = reinterpret_cast<float2*>(sWy + size_t(nL) * kYxJC);  // (g, jj, kYxRow)
  float2* sWx = sT2 + size_t(nG0) * kYxJC * kYxRow;                    // (jx, g), g < nG0p, row wxs
  int* sOff = reinterpret_cast<int*>(sWx + size_t(m0) * wxs);          // T1 row of each entry
  int* sF1 = sOff + nL;                                                // f1 of each entry
  int* sGs = sF1 + nL;                                                 // group starts
  int* sJx = sGs + nG0 + 1;                                            // kept rows, 64 per wave
  const float2* T1 = reinterpret_cast<const float2*>(static_cast<const float*>(sp.T1) + size_t(panel) * sp.t1PanelStride);
  for (int e = t; e < nL; e += 256) {
    sOff[e] = (sp.grpLine[e] * mz + jz) * ncol;
    sF1[e] = sp.grpF1[e];
  }
  for (int i = t; i <= nG0; i += 256) sGs[i] = sp.grpStart[i];
  for (int i = t; i < m0 * nG0p; i += 256) {
    const int jx = i / nG0p, g = i - jx * nG0p;
    float2 w = make_float2(0.f, 0.f);
    if (g < nG0) {
      const double2 d = p.tw0[(jx * p.g0val[g]) % m0];
      w = make_float2(float(d.x), float(d.y));
    }
    sWx[jx * wxs + g] = w;
  }
  __syncthreads();
  // T1 of the block: up to 8 loads per thread in flight before their LDS stores
  for (int i0 = t; i0 < nL * 16; i0 += 8 * 256) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + 256 * u, col = cb * 16 + (i & 15);
      v[u] = make_float2(0.f, 0.f);
      if (i < nL * 16 && col < ncol) v[u] = T1[size_t(sOff[i >> 4]) + col];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + 256 * u;
      if (i < nL * 16) sT1[(i >> 4) * kYxRow + (i & 15)] = v[u];
    }
  }
  const int nLinesYZ = p.m1 * p.m2;
  const int KS = (2 * nG0 + 3) / 4;  // <= 12, checked by the launcher
  const int c = cb * 16 + lr;
  for (int jy0 = 0; jy0 < my; jy0 += kYxJC) {
    const int nj = min(kYxJC, my - jy0);
    for (int i = t; i < nL * kYxJC; i += 256) {
      const int e = i / kYxJC, jj = i - e * kYxJC;
      sWy[i] = jj < nj ? sp.twy[((jy0 + jj) * sF1[e]) % my] : make_double2(0., 0.);
    }
    __syncthreads();  // also: T1 staged (first chunk)
    // thread (g, column): all kYxJC y coordinates of the chunk from one read of each T1 entry
    for (int i = t; i < nG0 * 16; i += 256) {
      const int cc = i & 15, g = i >> 4;
      double2 acc[kYxJC];
#pragma unroll
      for (int jj = 0; jj < kYxJC; ++jj) acc[jj] = make_double2(0., 0.);
      for (int e = sGs[g]; e < sGs[g + 1]; ++e) {
        const float2 xv = sT1[e * kYxRow + cc];
        const double2 x = make_double2(double(xv.x), double(xv.y));
#pragma unroll
        for (int jj = 0; jj < kYxJC; ++jj) {
          const double2 ww = sWy[e * kYxJC + jj];
          acc[jj].x += ww.x * x.x - ww.y * x.y;
          acc[jj].y += ww.x * x.y + ww.y * x.x;
        }
      }
#pragma unroll
      for (int jj = 0; jj < kYxJC; ++jj)
        sT2[(g * kYxJC + jj) * kYxRow + cc] = make_float2(float(acc[jj].x), float(acc[jj].y));
    }
    __syncthreads();
    for (int jw = wave; jw < nj; jw += 4) {
      const int slot = panel * nLinesYZ + jz * p.m1 + jy0 + jw;
      const unsigned long long mask = p.keepMask[slot];
      if (mask != 0ull) {
        int* jxs = sJx + wave * 64;
        wave_sync();  // the wave's previous slot has read its list
        if (lane < m0 && ((mask >> lane) & 1ull)) jxs[__popcll(mask & ((1ull << lane) - 1ull))] = lane;
        wave_sync();
        const int nkeep = __popcll(mask);
        float bq[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) {
          const int k = 4 * q + lg, g = k >> 1, kc = k & 1;
          float b = 0.f;
          if (q < KS && g < nG0) {
            const float2 v = sT2[(g * kYxJC + jw) * kYxRow + lr];
            b = kc ? v.y : v.x;
          }
          bq[q] = b;
        }
        float2* out = static_cast<float2*>(p.G3) + size_t(slot) * m0 * ncol + c;
        for (int m0r = 0; m0r < nkeep; m0r += 16) {
          const int nr = min(16, nkeep - m0r);
          const bool rowOk = lr < nr;
          const float2* wrow = sWx + (rowOk ? jxs[m0r + lr] : 0) * wxs;
          acc_t accr, acci;
#pragma unroll
          for (int i = 0; i < 4; ++i) accr[i] = acci[i] = 0.f;
#pragma unroll
          for (int q = 0; q < 12; ++q) {
            if (q < KS) {
              const int k = 4 * q + lg, g = k >> 1, kc = k & 1;
              const bool gOk = g < nG0;
              const float2 w = wrow[gOk ? g : nG0];
              const bool on = rowOk && gOk;
              const float a1 = on ? (kc ? -w.y : w.x) : 0.f;
              const float a2 = on ? (kc ? w.x : w.y) : 0.f;
              accr = Mfma4<float>::mac(a1, bq[q], accr);
              acci = Mfma4<float>::mac(a2, bq[q], acci);
            }
          }
          if (c < ncol) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = Mfma4<float>::row(lg, i);
              if (row < nr) out[size_t(jxs[m0r + row]) * ncol] = make_float2(accr[i], acci[i]);
            }
          }
        }
      }
    }
    __syncthreads();  // the next chunk rewrites the twiddles, T2 and the kept-row lists
  }
}

constexpr int kSepEpThreads = 64;  // k_sep_ep workgroup: one wave

// Row data of an rpp-row pass of k_sep_ep into LDS: U2 (row r at sU2 + r u2s,
// nw x nb2max), e2 (rpp x nb2max), omega and mode groups (rpp x nm) and the
// occupations (nCalc x rpp x nb2max).  The loads of the e2, omega and U2
// entries are issued before their LDS stores; rpp nm <= blockDim, rpp nb2max
// <= 2 blockDim.
__device__ inline void sep_ep_stage_rows(const CoupParams& p, const SepXmRow* rows, int m0r, int nr, int rpp,
                                         int nw, int nm, int u2s, double2* sU2, double* se2, double* sw3,
                                         int* sg3, double* sf2) {
  const int t = threadIdx.x, nb2max = p.nb2max, bs = blockDim.x;
  const int ne = rpp * nb2max, nwv = rpp * nm, nf = p.nCalc * ne;
  double ev[2] = {0., 0.}, wv = 0.;
  int gv = 0;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = t + bs * u;
    if (i < ne) {
      const int r = i / nb2max, b = i % nb2max;
      if (r < nr && b < rows[m0r + r].nb2) ev[u] = p.e2pool[rows[m0r + r].soff + b];
    }
  }
  if (t < nwv) {
    const int r = t / nm, nu = t % nm;
    if (r < nr) {
      wv = p.Wpool[size_t(rows[m0r + r].iq) * nm + nu];
      gv = p.qModeGrp[size_t(rows[m0r + r].iq) * nm + nu];
    }
  }
  const int per = nw * nb2max, nU = rpp * per;
  for (int i0 = t; i0 < nU; i0 += 4 * bs) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + bs * u;
      v[u] = make_double2(0., 0.);
      if (i < nU) {
        const int r = i / per, o = i % per;
        if (r < nr) {
          const SepXmRow& rr = rows[m0r + r];
          if (o < nw * rr.nb2) v[u] = p.U2pool[rr.u2off + o];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + bs * u;
      if (i < nU) sU2[(i / per) * u2s + i % per] = v[u];
    }
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
    if (t + bs * u < ne) se2[t + bs * u] = ev[u];
  if (t < nwv) {
    sw3[t] = wv;
    sg3[t] = gv;
  }
  for (int i = t; i < nf; i += bs) {
    const int c = i / ne, o = i % ne, r = o / nb2max, b = o % nb2max;
    double f = 0.;
    if (r < nr && b < rows[m0r + r].nb2) f = p.fermi[size_t(c) * p.numStates + rows[m0r + r].soff + b];
    sf2[i] = f;
  }
}

// epilogue of the rows k_sep_xg wrote: the electron rotation from each
// thread's g3 slice (registers), then the phonon data and sep_epilogue_tail;
// nm threads per row, passes of epRows rows (whole rows per wave, nm <= NMX
// <= 16).  The slice and the eigenvector column are never live together.  CM calculations
// and GM members of the b1 group are held per thread: every panel of the
// launch has gsize <= GM (checked on the host).
template <typename T, int CM, int NMX, int GM>
__global__ void __launch_bounds__(kSepEpThreads, 2) k_sep_ep(CoupParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  constexpr int NWX = 16;
  const int nLinesYZ = p.m1 * p.m2;
  const int nTilesTot = p.nPanels * nLinesYZ;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int Lx = (bid % 8) * perXcd + bid / 8;
  if (Lx >= nTilesTot) return;
  const int yz = Lx / p.nPanels, panel = Lx % p.nPanels;
  const int slot = panel * nLinesYZ + yz;
  const int jy = yz % p.m1, jz = yz / p.m1;
  const int t = threadIdx.x;
  const unsigned long long mask = p.keepMask[slot];
  if (mask == 0ull) {
    if (t < kGMax * kCMax) p.partial[size_t(slot) * kGMax * kCMax + t] = 0.;
    return;
  }
  const int nm = p.nm, nw = p.nw, ncol = nm * nw;
  const int kk = p.panelKK[panel];
  const int b1p = p.panelB1[panel];
  const int kkMesh = p.kkMesh[kk];
  const int u2s = nw * p.nb2max + 1, gps = nm * p.nb2max + 1;
  const int rpp = p.epRows, rpw = rpp / (kSepEpThreads / 64);
  SepXmRow* rows = reinterpret_cast<SepXmRow*>(smem);                          // 64
  double2* sU2 = reinterpret_cast<double2*>(smem + (64 * sizeof(SepXmRow) + 15) / 16 * 16);  // rpp x u2s
  double2* gp = sU2 + size_t(rpp) * u2s;                                       // rpp x gps
  double* se2 = reinterpret_cast<double*>(gp + size_t(rpp) * gps);             // rpp x nb2max
  double* sw3 = se2 + size_t(rpp) * p.nb2max;                                  // rpp x nm
  double* sf2 = sw3 + size_t(rpp) * nm;                                        // nCalc x rpp x nb2max
  double* se1 = sf2 + size_t(p.nCalc) * rpp * p.nb2max;                        // kGMax
  double* gsq = se1 + kGMax;                                                   // rpp x nm x nb2max
  int* sg3 = reinterpret_cast<int*>(gsq + size_t(rpp) * nm * p.nb2max);        // rpp x nm
  if (t < p.panelGSize[panel]) se1[t] = p.e1pool[p.kkStateOff[kk] + p.panelGStart[panel] + t];
  const int nkeep = __popcll(mask);
  if (t < p.m0 && ((mask >> t) & 1ull)) {
    const int mesh = (jz * p.m1 + jy) * p.m0 + t;
    const int4 mr = p.meshRow[mesh];
    rows[__popcll(mask & ((1ull << t) - 1ull))] =
        SepXmRow{t, mr.x, mesh_diff_index(mesh, kkMesh, p.m0, p.m1, p.m2), mr.y, mr.z, mr.w};
  }
  if (t == 0 && p.keptRows) atomicAdd(p.keptRows, (unsigned long long)nkeep);
  __syncthreads();
  if (t == 0 && p.keptBands) {
    unsigned long long nb = 0;
    for (int i = 0; i < nkeep; ++i) nb += rows[i].nb2;
    atomicAdd(p.keptBands, nb);
  }
  double part[GM][CM];
#pragma unroll
  for (int gi = 0; gi < GM; ++gi)
#pragma unroll
    for (int c = 0; c < CM; ++c) part[gi][c] = 0.;
  const int gstart = p.panelGStart[panel], gsize = p.panelGSize[panel];
  const double* e1 = p.e1pool + p.kkStateOff[kk] + gstart;
  const double inv_g = 1. / double(float(gsize));
  const T2v* G3 = static_cast<const T2v*>(p.G3) + size_t(slot) * p.m0 * ncol;
  // rpw whole rows of nm threads in each wave; the lanes beyond them idle
  const int wave = t >> 6, lane = t & 63;
  const bool laneOn = lane < rpw * nm;
  const int r = laneOn ? wave * rpw + lane / nm : rpp, sub = laneOn ? lane % nm : 0, rl = r;
  for (int m0r = 0; m0r < nkeep; m0r += rpp) {
    const int nr = min(rpp, nkeep - m0r);
    const bool rowOn = r < nr;
    const bool modeOn = rowOn && sub < nm;
    // the thread's g3 slice (mode sub of its row), loads issued before the
    // row staging's
    T2v g[NWX];
    if (modeOn) {
      const T2v* src = G3 + size_t(rows[m0r + r].jx) * ncol + sub * nw;
      if (sizeof(T) == 4 && nw == NWX) {
        // full slice: 16-byte loads (src is a multiple of 16 elements from
        // the table's start)
        const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
        for (int h = 0; h < NWX / 2; ++h) {
          const float4 v = s4[h];
          g[2 * h].x = v.x;
          g[2 * h].y = v.y;
          g[2 * h + 1].x = v.z;
          g[2 * h + 1].y = v.w;
        }
      } else {
#pragma unroll
        for (int w2 = 0; w2 < NWX; ++w2)
          if (w2 < nw) g[w2] = src[w2];
      }
    }
    sep_ep_stage_rows(p, rows, m0r, nr, rpp, nw, nm, u2s, sU2, se2, sw3, sg3, sf2);
    __syncthreads();
    const SepXmRow ri = rows[m0r + (rowOn ? r : 0)];
    const int nb2 = rowOn ? ri.nb2 : 0;
    const double w3 = modeOn ? sw3[r * nm + sub] : 0.;
    // g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2), thread nu = sub
    if (modeOn) {
      for (int b2 = 0; b2 < nb2; ++b2) {
        const double2* u = sU2 + size_t(r) * u2s + b2 * nw;
        double sx = 0., sy = 0.;
#pragma unroll
        for (int w2 = 0; w2 < NWX; ++w2)
          if (w2 < nw) {
            const double xr = double(g[w2].x), xi = double(g[w2].y);
            sx += u[w2].x * xr + u[w2].y * xi;
            sy += u[w2].x * xi - u[w2].y * xr;
          }
        gp[rl * gps + sub * p.nb2max + b2] = make_double2(sx, sy);
      }
    }
    // phonon eigenvector column and Bose factors of (row, mode sub)
    T2v e3c[NMX];
    double n3[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) n3[c] = 0.;
    if (modeOn) {
      const T2v* e3 = static_cast<const T2v*>(p.E3T) + size_t(ri.iq) * nm * nm + sub;
#pragma unroll
      for (int nu = 0; nu < NMX; ++nu)
        if (nu < nm) e3c[nu] = e3[nu * nm];
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (c < p.nCalc) n3[c] = p.bose[(size_t(ri.iq) * nm + sub) * p.nCalc + c];
    }
    wave_sync();
    sep_epilogue_tail_lds<T, NMX, CM, GM>(p, ri, r, rl, sub, nb2, modeOn, w3, gp, gps, gsq, se2, sg3, sf2, se1, e3c,
                                      n3, part, e1, gsize, inv_g, rpp);
    __syncthreads();  // the next tile rewrites the row data
  }
  reduce_partials<CM, GM, kSepEpThreads / 64>(p, slot, smem, part);
}

inline size_t sep_ep_lds_bytes(int nw, int nm, int nb2max, int nCalc, int rpp) {
  size_t b = (64 * sizeof(SepXmRow) + 15) / 16 * 16;
  b += size_t(rpp) * (nw * nb2max + 1) * sizeof(double2) + size_t(rpp) * (nm * nb2max + 1) * sizeof(double2) +
       size_t(rpp) * nb2max * sizeof(double) + size_t(rpp) * nm * sizeof(double) +
       size_t(nCalc) * rpp * nb2max * sizeof(double) + size_t(kGMax) * sizeof(double) +
       size_t(rpp) * nm * nb2max * sizeof(double) + size_t(rpp) * nm * sizeof(int);
  return std::max(b, size_t(kGMax) * kCMax * 4 * sizeof(double)) + 64;
}

// ------------------------------------------------------------------ split epilogue, lean form
// Only what the epilogue of the split last stage reads (fixed Gaussian
// smearing).  A small block: the kernel keeps its scalars in SGPRs instead of
// spilling the full CoupParams' into VGPR lanes.
struct EpParams {
  const unsigned long long* keepMask;
  double* partial;  // (slot, kGMax, kCMax)
  const int* panelKK;
  const int* panelGStart;
  const int* panelGSize;
  const int* kkMesh;
  const int* kkStateOff;
  const int4* meshRow;     // per mesh point: (row, nb2, u2off, stateOff)
  const double* e1pool;
  const double* e2pool;
  const double* Wpool;     // (q, nu)
  const int* qModeGrp;     // (q, nu): first mode of the group | size << 8
  const double2* U2pool;
  const double* fermi;     // (calc, state)
  const double* bose;      // (q, nu, calc)
  const void* E3T;         // (q, nu, nu2) in the transform's precision
  const void* G3;          // (slot, jx, (nu, w2)) in the transform's precision; with the
                           // rotation done by the x-stage, (slot, jx, b2, nu) complex FP64
  const double* screenOf;  // per outer state: screen width of the panel its b1 group starts (null: screen)
  unsigned long long* keptRows;
  unsigned long long* keptBands;
  const unsigned char* slotOn;  // compute only slots with slotOn != 0 (null: all)
  const int* gate;              // run only when *gate != 0 (null: always)
  int slotZero;                 // slots off: 1 = write zero partials, 0 = leave them
  double invSigma, deltaPref, norm, cutoff, screen;
  int m0, m1, m2, nPanels, nm, nw, nCalc, numStates;
  int nb2max;              // g' rows of the rotating x-stage hold nb2max nm entries
  int rpw;                 // rows per pass: whole rows of nm lanes, rpw nm <= 64
};

// Epilogue of the rows k_sep_xg wrote, one wave per slot (panel, yz line),
// passes of rpw rows with lane (row r, mode sub).  The same operations in the
// same order as k_sep_ep (so the same linewidths), arranged so that little is
// live at once:
//   A  electron rotation g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2),
//      nu = sub, each g3 value converted once (outer loop over w2, NB2
//      accumulators), U2 of the pass in LDS;
//   B  phonon rotation g(nu2, b2) = sum_nu e3(nu, nu2) g'(nu, b2), nu2 = sub,
//      the eigenvector column in registers, then |g|^2 and symmetrizeCoupling
//      along b2 (scattering.cpp:1250-1284) on the lane's own entries;
//   C  symmetrizeCoupling along nu (:1286-1316) from the group's entries and
//      the Gaussian-delta weights of el_scattering.cpp:277-380 into per-lane
//      partial sums (CM calculations, GM members of the b1 group).
// Phases A and B run the pass's largest band count for every row (uniform
// branches; a row's U2 columns beyond its own count are zero).  NB2 >= nb2max
// is the LDS stride of a row's bands, NMX >= nm the eigenvector column length.
// Phase A of k_sep_ep2 for NB bands: g'(nu, b2) = sum_w2 conj(U2(w2, b2)) g3(nu, w2)
// for the lane's mode, w2 in order.  No branches inside: entries beyond nw
// are zero in g and in U2, so every term adds an exact zero.
template <typename T2v, int NWX, int NB2, int NB>
__device__ __forceinline__ void ep2_rot_el(const T2v (&g)[NWX], const double2* u, double2* out) {
  double ar[NB], ai[NB];
#pragma unroll
  for (int b2 = 0; b2 < NB; ++b2) ar[b2] = ai[b2] = 0.;
#pragma unroll
  for (int w2 = 0; w2 < NWX; ++w2) {
    const double xr = double(g[w2].x), xi = double(g[w2].y);
#pragma unroll
    for (int b2 = 0; b2 < NB; ++b2) {
      const double2 c = u[b2 * NWX + w2];
      ar[b2] += c.x * xr + c.y * xi;
      ai[b2] += c.x * xi - c.y * xr;
    }
  }
#pragma unroll
  for (int b2 = 0; b2 < NB; ++b2) out[b2] = make_double2(ar[b2], ai[b2]);
}

// Phase B of k_sep_ep2 for NB bands: g(nu2, b2) = sum_nu e3(nu, nu2) g'(nu, b2)
// for the lane's mode nu2, nu in order, and |g|^2 into own.  Modes beyond nm
// have zero eigenvector entries and read the last mode's (finite) g'.
template <typename T2v, int NMX, int NB2, int NB>
__device__ __forceinline__ void ep2_rot_ph(const T2v (&e3c)[NMX], const double2* x, int nm, double* own) {
  double gr[NB], gi[NB];
#pragma unroll
  for (int b2 = 0; b2 < NB; ++b2) gr[b2] = gi[b2] = 0.;
#pragma unroll
  for (int nu = 0; nu < NMX; ++nu) {
    const double er = double(e3c[nu].x), ei = double(e3c[nu].y);
    const double2* xn = x + min(nu, nm - 1) * NB2;
#pragma unroll
    for (int b2 = 0; b2 < NB; ++b2) {
      const double2 v = xn[b2];
      gr[b2] += v.x * er - v.y * ei;
      gi[b2] += v.x * ei + v.y * er;
    }
  }
#pragma unroll
  for (int b2 = 0; b2 < NB; ++b2) own[b2] = gr[b2] * gr[b2] + gi[b2] * gi[b2];
}

template <typename T, int CM, int NMX, int GM, int NB2, bool PRE = false>
__global__ void __launch_bounds__(64, 3) k_sep_ep2(EpParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename Vec2<T>::type T2v;
  constexpr int NWX = 16;
  constexpr int U2S = NWX * NB2 + 1;  // double2 per row of the U2 table (odd: rows in different banks)
  constexpr int RMX = NB2 <= 4 ? 5 : 4;           // rows per pass at most (the launcher's cap)
  constexpr int UPL = (NWX * NB2 * RMX + 63) / 64;  // U2 entries per lane
  const int nLinesYZ = p.m1 * p.m2;
  const int nTilesTot = p.nPanels * nLinesYZ;
  const int bid = blockIdx.x;
  const int perXcd = gridDim.x / 8;
  const int Lx = (bid % 8) * perXcd + bid / 8;
  if (Lx >= nTilesTot) return;
  const int yz = Lx / p.nPanels, panel = Lx % p.nPanels;
  const int slot = panel * nLinesYZ + yz;
  const int jy = yz % p.m1, jz = yz / p.m1;
  const int t = threadIdx.x;
  if (p.gate && *p.gate == 0) return;
  if (p.slotOn && !p.slotOn[slot]) {
    if (p.slotZero && t < kGMax * kCMax) p.partial[size_t(slot) * kGMax * kCMax + t] = 0.;
    return;
  }
  const unsigned long long mask = p.keepMask[slot];
  if (mask == 0ull) {
    if (t < kGMax * kCMax) p.partial[size_t(slot) * kGMax * kCMax + t] = 0.;
    return;
  }
  const int nm = p.nm, nw = p.nw, ncol = nm * nw, R = p.rpw;
  const int kk = p.panelKK[panel];
  const int gstart = p.panelGStart[panel], gsize = __builtin_amdgcn_readfirstlane(p.panelGSize[panel]);
  const int is1 = p.kkStateOff[kk] + gstart;
  const double scrW = p.screen > 0. ? (p.screenOf ? p.screenOf[is1] : p.screen) : 0.;
  const double inv_g = 1. / double(float(gsize));
  // LDS: rows (64) | U2 (R x U2S double2), later |g|^2 (R x nm x NB2) | g' (R x nm x NB2 + 1 per row)
  //      | e2 (R x NB2) | occupations (CM x R x NB2) | e1 (kGMax)
  SepXmRow* rows = reinterpret_cast<SepXmRow*>(smem);
  double2* sU2 = reinterpret_cast<double2*>(smem + (64 * sizeof(SepXmRow) + 15) / 16 * 16);
  double* gsq = reinterpret_cast<double*>(sU2);
  const int gps = nm * NB2 + 1;
  double2* gp = sU2 + size_t(R) * U2S;
  double* se2 = reinterpret_cast<double*>(gp + size_t(R) * gps);
  double* sf2 = se2 + R * NB2;
  double* se1 = sf2 + CM * R * NB2;
  if (t < gsize) se1[t] = p.e1pool[is1 + t];
  const int nkeep = __popcll(mask);
  {
    const int kkMesh = p.kkMesh[kk];
    if (t < p.m0 && ((mask >> t) & 1ull)) {
      const int mesh = (jz * p.m1 + jy) * p.m0 + t;
      const int4 mr = p.meshRow[mesh];
      rows[__popcll(mask & ((1ull << t) - 1ull))] =
          SepXmRow{t, mr.x, mesh_diff_index(mesh, kkMesh, p.m0, p.m1, p.m2), mr.y, mr.z, mr.w};
    }
  }
  __syncthreads();
  if (t == 0) {
    if (p.keptRows) atomicAdd(p.keptRows, (unsigned long long)nkeep);
    if (p.keptBands) {
      unsigned long long nb = 0;
      for (int i = 0; i < nkeep; ++i) nb += rows[i].nb2;
      atomicAdd(p.keptBands, nb);
    }
  }
  // lane (r, sub): r < R rows of a pass, whole rows of nm lanes; the rest idle
  const bool laneOn = t < R * nm;
  const int r = laneOn ? t / nm : 0, sub = laneOn ? t % nm : 0;
  const T2v* G3 = static_cast<const T2v*>(p.G3) + size_t(slot) * p.m0 * ncol;
  const int rowG = p.nb2max * nm;  // PRE: g' row length, (b2, nu)
  const double2* GP = static_cast<const double2*>(p.G3) + size_t(slot) * p.m0 * rowG;
  const T2v* E3T = static_cast<const T2v*>(p.E3T);
  double part[GM][CM];
#pragma unroll
  for (int gi = 0; gi < GM; ++gi)
#pragma unroll
    for (int c = 0; c < CM; ++c) part[gi][c] = 0.;

  // registers of one pass, loaded at its start
  T2v g[NWX];         // g3(nu = sub, w2) of the lane's row
  double2 gv[NB2];    // PRE: g'(nu = sub, b2) of the lane's row
  T2v e3c[NMX];       // e3(nu, nu2 = sub) at the row's q
  double2 u2v[UPL];   // U2 entries of the pass (lane-strided over R x NWX x NB2)
  double e2v = 0., f2v[CM], w3 = 0., n3[CM];
  int grp = 0;
  auto load_pass = [&](int m0r) {
    const int nr = min(R, nkeep - m0r);
    const bool rowOn = laneOn && r < nr;
    const SepXmRow ri = rows[m0r + (rowOn ? r : 0)];
    if (rowOn) {
      if (PRE) {
#pragma unroll
        for (int b2 = 0; b2 < NB2; ++b2)
          gv[b2] = b2 < ri.nb2 ? GP[size_t(ri.jx) * rowG + b2 * nm + sub] : make_double2(0., 0.);
      } else if (sizeof(T) == 4 && nw == NWX) {
        const float4* s4 = reinterpret_cast<const float4*>(G3 + size_t(ri.jx) * ncol + sub * nw);
#pragma unroll
        for (int h = 0; h < NWX / 2; ++h) {
          const float4 v = s4[h];
          g[2 * h].x = v.x;
          g[2 * h].y = v.y;
          g[2 * h + 1].x = v.z;
          g[2 * h + 1].y = v.w;
        }
      } else {
        const T2v* src = G3 + size_t(ri.jx) * ncol + sub * nw;
#pragma unroll
        for (int w2 = 0; w2 < NWX; ++w2) g[w2] = w2 < nw ? src[w2] : T2v{T(0), T(0)};
      }
      const T2v* e3 = E3T + size_t(ri.iq) * nm * nm + sub;
#pragma unroll
      for (int nu = 0; nu < NMX; ++nu) e3c[nu] = nu < nm ? e3[nu * nm] : T2v{T(0), T(0)};
      w3 = p.Wpool[size_t(ri.iq) * nm + sub];
      grp = p.qModeGrp[size_t(ri.iq) * nm + sub];
#pragma unroll
      for (int c = 0; c < CM; ++c) n3[c] = c < p.nCalc ? p.bose[(size_t(ri.iq) * nm + sub) * p.nCalc + c] : 0.;
    } else {
      w3 = 0.;
      grp = 0;
    }
    // U2 of the pass: entry i = (row, b2, w2), w2 fastest, NWX x NB2 per row
#pragma unroll
    for (int u = 0; u < (PRE ? 0 : UPL); ++u) {
      const int i = t + 64 * u, rr = i / (NWX * NB2), o = i % (NWX * NB2), b2 = o / NWX, w2 = o % NWX;
      double2 v = make_double2(0., 0.);
      if (rr < nr) {
        const SepXmRow& q = rows[m0r + rr];
        if (b2 < q.nb2 && w2 < nw) v = p.U2pool[q.u2off + b2 * nw + w2];
      }
      u2v[u] = v;
    }
    // band energies and occupations: lane = (row, b2) for lanes < R NB2
    e2v = 0.;
#pragma unroll
    for (int c = 0; c < CM; ++c) f2v[c] = 0.;
    if (t < R * NB2) {
      const int rr = t / NB2, b2 = t % NB2;
      if (rr < nr && b2 < rows[m0r + rr].nb2) {
        const int st = rows[m0r + rr].soff + b2;
        e2v = p.e2pool[st];
#pragma unroll
        for (int c = 0; c < CM; ++c)
          if (c < p.nCalc) f2v[c] = p.fermi[size_t(c) * p.numStates + st];
      }
    }
  };
  for (int m0r = 0; m0r < nkeep; m0r += R) {
    load_pass(m0r);
    const int nr = min(R, nkeep - m0r);
    const bool rowOn = laneOn && r < nr;
    const int nb2 = rowOn ? rows[m0r + r].nb2 : 0;
    // largest band count of the pass's rows, uniform over the wave: phases A
    // and B run that many bands for every row (a row's U2 columns beyond its
    // own count are zero, so those entries come out zero and are never read)
    int nbw = 0;
    for (int i = 0; i < nr; ++i) nbw = max(nbw, rows[m0r + i].nb2);
    nbw = __builtin_amdgcn_readfirstlane(nbw);
    // row data of the pass into LDS
#pragma unroll
    for (int u = 0; u < (PRE ? 0 : UPL); ++u) {
      const int i = t + 64 * u, rr = i / (NWX * NB2);
      if (rr < R) sU2[rr * U2S + i % (NWX * NB2)] = u2v[u];  // R <= RMX
    }
    if (t < R * NB2) {
      se2[t] = e2v;
#pragma unroll
      for (int c = 0; c < CM; ++c) sf2[c * R * NB2 + t] = f2v[c];
    }
    __syncthreads();
    // A: electron rotation, thread nu = sub (PRE: done by the x-stage kernel)
    if (PRE && rowOn) {
      double2* out = gp + r * gps + sub * NB2;
#pragma unroll
      for (int b2 = 0; b2 < NB2; ++b2) out[b2] = gv[b2];
    }
    if (!PRE && rowOn) {
      const double2* u = sU2 + r * U2S;
      double2* out = gp + r * gps + sub * NB2;
      switch (nbw) {
        case 1: ep2_rot_el<T2v, NWX, NB2, 1>(g, u, out); break;
        case 2: ep2_rot_el<T2v, NWX, NB2, 2>(g, u, out); break;
        case 3: ep2_rot_el<T2v, NWX, NB2, 3>(g, u, out); break;
        case 4: ep2_rot_el<T2v, NWX, NB2, 4>(g, u, out); break;
        default:
          if constexpr (NB2 > 4) ep2_rot_el<T2v, NWX, NB2, NB2>(g, u, out);
      }
    }
    wave_sync();
    // B: phonon rotation, thread nu2 = sub, then |g|^2 and the b2 groups
    double* own = gsq + (size_t(r) * nm + sub) * NB2;
    if (rowOn) {
      const double2* x = gp + r * gps;
      switch (nbw) {
        case 1: ep2_rot_ph<T2v, NMX, NB2, 1>(e3c, x, nm, own); break;
        case 2: ep2_rot_ph<T2v, NMX, NB2, 2>(e3c, x, nm, own); break;
        case 3: ep2_rot_ph<T2v, NMX, NB2, 3>(e3c, x, nm, own); break;
        case 4: ep2_rot_ph<T2v, NMX, NB2, 4>(e3c, x, nm, own); break;
        default:
          if constexpr (NB2 > 4) ep2_rot_ph<T2v, NMX, NB2, NB2>(e3c, x, nm, own);
      }
      const double* e2r = se2 + r * NB2;
      for (int b = 0; b < nb2;) {
        int deg = 0;
        for (int i = b; i < nb2; ++i) {
          if (fabs(e2r[b] - e2r[i]) < 1.0e-6) deg++;
          else break;
        }
        if (deg > 1) {
          double s2 = 0.;
          for (int i = 0; i < deg; ++i) s2 += own[b + i] / double(float(deg));
          for (int i = 0; i < deg; ++i) own[b + i] = s2;
        }
        b += deg;
      }
    }
    wave_sync();
    const double w3c = w3;
    const int grpc = grp;
    double n3c[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) n3c[c] = n3[c];
    // C: the nu groups and the Gaussian weights
    if (rowOn && w3c >= p.cutoff) {
      const int mystart = grpc & 0xff, mydeg = grpc >> 8;
      const double* g0 = gsq + (size_t(r) * nm + mystart) * NB2;
      const double wfac = p.norm / w3c * 3.14159265358979323846;  // once per lane, not per term
#pragma unroll 1
      for (int b2 = 0; b2 < nb2; ++b2) {
        double vb;
        if (mydeg > 1) {
          vb = 0.;
          for (int i = 0; i < mydeg; ++i) vb += g0[size_t(i) * NB2 + b2] / double(float(mydeg));
        } else {
          vb = own[b2];
        }
        const double en2 = se2[r * NB2 + b2];
        const double cpl = vb * inv_g;
#pragma unroll
        for (int gi = 0; gi < GM; ++gi) {
          if (gi >= gsize) break;
          const double x1 = se1[gi] - en2 + w3c, x2 = se1[gi] - en2 - w3c;
          if (p.screen > 0. && fmin(fabs(x1), fabs(x2)) * p.invSigma >= scrW) continue;
          const double d1 = gaussian_delta(x1, p.invSigma, p.deltaPref);
          const double d2 = gaussian_delta(x2, p.invSigma, p.deltaPref);
          if (d1 <= 0. && d2 <= 0.) continue;
#pragma unroll
          for (int c = 0; c < CM; ++c) {
            if (c >= p.nCalc) break;
            const double f2 = sf2[(c * R + r) * NB2 + b2];
            part[gi][c] += cpl * ((f2 + n3c[c]) * d1 + (1. - f2 + n3c[c]) * d2) * wfac;
          }
        }
      }
    }
    __syncthreads();  // the next pass rewrites the row data
  }
  // wave reduction of the partial linewidths into partial[slot]
  double* red = reinterpret_cast<double*>(smem);
  const int lane = t & 63;
#pragma unroll
  for (int gi = 0; gi < GM; ++gi)
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      double v = part[gi][c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[gi * kCMax + c] = v;
    }
  __syncthreads();
  if (t < kGMax * kCMax) {
    const int c = t % kCMax, gi = t / kCMax;
    p.partial[size_t(slot) * kGMax * kCMax + t] = (c < CM && gi < GM) ? red[t] : 0.;
  }
}

inline size_t sep_ep2_lds_bytes(int nm, int nb2, int cm, int rpw) {
  size_t b = (64 * sizeof(SepXmRow) + 15) / 16 * 16;
  b += size_t(rpw) * (16 * nb2 + 1) * sizeof(double2);              // U2, later |g|^2
  b += size_t(rpw) * (nm * nb2 + 1) * sizeof(double2);              // g'
  b += size_t(rpw) * nb2 * sizeof(double) * (1 + cm);               // e2, occupations
  b += size_t(kGMax) * sizeof(double);                              // e1
  return std::max(b, size_t(kGMax) * kCMax * sizeof(double)) + 64;
}

constexpr int kXminBlocks = 64;

// Closest approach to resonance of every panel of the chunk, for the energy
// screen widths: the minimum over the panel's rows (inner points of the
// range), their bands b2, the modes above the cutoff and the b1 group's
// energies e1 of |x| sigma, x = (e1 - e2 +- omega) / sigma, over the pairs
// with x <= 6 (the others have zero weight under the one-sided cut).  Grid (kXminBlocks, panels): block b
// takes mesh points b * 256 + t, strided by 256 * kXminBlocks.  A block first
// evaluates only the rows that pass the interval test at the configured
// width n: a row that fails it is at least n sigma from resonance in every
// pair, so once the block has found a pair closer than n the skipped rows
// cannot hold its minimum.  Otherwise it evaluates every row.  One partial
// minimum (energy units, HUGE_VAL when there is no pair) per block.
static __global__ void __launch_bounds__(256) k_sep_xmin(CoupParams p) {
  const int panel = blockIdx.y, t = threadIdx.x;
  const int kk = p.panelKK[panel];
  const int kmesh = p.kkMesh[kk];
  const double* e1 = p.e1pool + p.kkStateOff[kk] + p.panelGStart[panel];
  const int gsz = p.panelGSize[panel];
  const double lim = p.screen / p.invSigma;
  const int nmesh = p.m0 * p.m1 * p.m2;
  __shared__ double red[256];
  double best = HUGE_VAL, bmin = HUGE_VAL;
  for (int pass = 0; pass < 2; ++pass) {
    for (int mesh = blockIdx.x * 256 + t; mesh < nmesh; mesh += 256 * kXminBlocks) {
      const int row = p.meshToRow[mesh];
      if (row < p.rowLo || row >= p.rowHi) continue;
      const int iq = mesh_diff_index(mesh, kmesh, p.m0, p.m1, p.m2);
      const double wlo = p.qWlo[iq], whi = p.qWhi[iq];
      if (whi < wlo) continue;  // no mode above the cutoff
      if (pass == 0) {
        const double2 er = p.rowE2range[row];
        bool keep = false;
        for (int gi = 0; gi < gsz; ++gi) {
          const double a_lo = e1[gi] - er.y + wlo, a_hi = e1[gi] - er.x + whi;
          const double b_lo = e1[gi] - er.y - whi, b_hi = e1[gi] - er.x - wlo;
          keep |= (a_hi > -lim && a_lo < lim) || (b_hi > -lim && b_lo < lim);
        }
        if (!keep) continue;
      }
      const double* w = p.Wpool + size_t(iq) * p.nm;
      const int soff = p.rowStateOff[row], nb2 = p.rowNb2[row];
      for (int b2 = 0; b2 < nb2; ++b2) {
        const double e2 = p.e2pool[soff + b2];
        for (int gi = 0; gi < gsz; ++gi) {
          const double d = e1[gi] - e2;
          for (int nu = 0; nu < p.nm; ++nu) {
            const double om = w[nu];
            if (om < p.cutoff) continue;
            // only pairs the one-sided Gaussian can weigh (x <= 6, delta_function.cpp:43)
            const double a = d + om, b = d - om;
            if (a * p.invSigma <= 6.) best = fmin(best, fabs(a));
            if (b * p.invSigma <= 6.) best = fmin(best, fabs(b));
          }
        }
      }
    }
    red[t] = best;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) red[t] = fmin(red[t], red[t + s]);
      __syncthreads();
    }
    bmin = red[0];
    __syncthreads();
    if (bmin < lim) break;  // block-uniform
  }
  if (t == 0) p.xminPart[size_t(panel) * kXminBlocks + blockIdx.x] = bmin;
}

// Screen width of each panel: sqrt(xmin^2 + n^2) sigma, with xmin its closest
// approach to resonance in units of sigma.  A pair that the screen drops then
// weighs at most exp(-n^2) of the panel's most resonant weighed pair, whatever
// the panel's distance to resonance (a panel whose pairs all lie in the
// Gaussian tails keeps its closest rows).  One thread per panel.
static __global__ void __launch_bounds__(64) k_sep_xmin_fin(CoupParams p) {
  const int panel = blockIdx.x * 64 + threadIdx.x;
  if (panel >= p.nPanels) return;
  double m = HUGE_VAL;
  for (int b = 0; b < kXminBlocks; ++b) m = fmin(m, p.xminPart[size_t(panel) * kXminBlocks + b]);
  const double x = m * p.invSigma;
  const int is1 = p.kkStateOff[p.panelKK[panel]] + p.panelGStart[panel];
  p.screenOf[is1] = x < HUGE_VAL ? sqrt(x * x + p.screen * p.screen) : p.screen;
}

// Row lists of the last stage: for every (panel, yz) slot, bit jx of
// keepMask[slot] is set when mesh point (jx, jy, jz) is an inner point of the
// range and passes the energy screen (same interval test as k_sep_x; every
// valid row when the screen is off).  One thread per slot, m0 <= 64.
static __global__ void __launch_bounds__(256) k_sep_screen(CoupParams p) {
  const int nLinesYZ = p.m1 * p.m2;
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= p.nPanels * nLinesYZ) return;
  const int panel = slot / nLinesYZ, yz = slot % nLinesYZ;
  const int jy = yz % p.m1, jz = yz / p.m1;
  const int kk = p.panelKK[panel];
  const int kmesh = p.kkMesh[kk];
  const double* e1 = p.e1pool + p.kkStateOff[kk] + p.panelGStart[panel];
  const int gsz = p.panelGSize[panel];
  const double lim = screen_width(p, int(e1 - p.e1pool)) / p.invSigma;
  unsigned long long mask = 0ull;
  for (int jx = 0; jx < p.m0; ++jx) {
    const int mesh = (jz * p.m1 + jy) * p.m0 + jx;
    const int row = p.meshRow[mesh].x;
    if (row < p.rowLo || row >= p.rowHi) continue;
    bool keep = true;
    if (p.screen > 0.) {
      const int iq = mesh_diff_index(mesh, kmesh, p.m0, p.m1, p.m2);
      const double wlo = p.qWlo[iq], whi = p.qWhi[iq];
      const double2 er = p.rowE2range[row];
      keep = false;
      if (whi >= wlo)  // at least one mode above the cutoff
        for (int gi = 0; gi < gsz; ++gi) {
          const double a_lo = e1[gi] - er.y + wlo, a_hi = e1[gi] - er.x + whi;
          const double b_lo = e1[gi] - er.y - whi, b_hi = e1[gi] - er.x - wlo;
          keep |= (a_hi > -lim && a_lo < lim) || (b_hi > -lim && b_lo < lim);
        }
    }
    if (keep) mask |= 1ull << jx;
  }
  p.keepMask[slot] = mask;
}

inline size_t sep_xm_lds_bytes(int elem, int nG0, int m0, int nw, int nm, int nb2max, int rpp, int nCalc) {
  auto up = [](size_t b) { return (b + 15) / 16 * 16; };
  size_t b = up(size_t(m0) * sizeof(SepXmRow) + (8 + nG0) * sizeof(int));
  b = up(b + size_t(m0) * 2 * elem + size_t(m0) * (nG0 + 1) * 2 * elem);
  const size_t g3 = std::max(size_t(16) * nm * (2 * nw + 1) * elem, size_t(rpp) * nm * nb2max * sizeof(double));
  b += up(g3);
  b += size_t(16) * (nw * nb2max + 1) * sizeof(double2) + size_t(rpp) * (nm * nb2max + 1) * sizeof(double2) +
       size_t(16) * nb2max * sizeof(double) + size_t(16) * nm * sizeof(double) +
       size_t(nCalc) * 16 * nb2max * sizeof(double) + size_t(kGMax) * sizeof(double) +
       size_t(16) * nm * sizeof(int);
  return std::max(b, size_t(kGMax) * kCMax * 4 * sizeof(double)) + 64;
}

// maxLineE / maxGrpE: largest number of entries of a line / group (LDS sizing).
template <typename T>
void launch_separable_zy(const SepParams& sp, int maxLineE, int maxGrpE, hipStream_t s, bool yStage = true) {
  if (sp.nPanels <= 0) return;
  const size_t ldsZ = sep_stage_lds_bytes(sizeof(T), maxLineE, sp.ncol, sp.mz);
  const size_t ldsY = sep_stage_lds_bytes(sizeof(T), maxGrpE, sp.ncol, sp.my);
  PHX_REQUIRE(ldsZ <= 160 * 1024 && ldsY <= 160 * 1024, "separable transform: LDS budget exceeded");
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_sep_z<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(std::max<size_t>(ldsZ, 16))));
  PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_sep_y<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(std::max<size_t>(ldsY, 16))));
  hipLaunchKernelGGL(k_sep_z<T>, dim3(sp.nLines, sp.nPanels), dim3(256), ldsZ, s, sp);
  PHX_HIP(hipGetLastError());
  if (!yStage) return;  // the x-stage kernel runs the y-stage itself (k_sep_yx)
  hipLaunchKernelGGL(k_sep_y<T>, dim3(sp.nG0, sp.mz, sp.nPanels), dim3(256), ldsY, s, sp);
  PHX_HIP(hipGetLastError());
}

// Where the split last stage's epilogue kernel runs: on the launch's stream
// (ep null), or on the stream ep once the x-stage kernel is done, so that it
// overlaps the work enqueued on the launch's stream after it.
struct EpFork {
  hipStream_t ep = nullptr;
  hipEvent_t xgDone = nullptr;   // recorded on the launch's stream after the x-stage kernel (ep set)
  hipEvent_t epStart = nullptr;  // timing: recorded on the epilogue's stream before it (null: off)
  hipEvent_t xgStart = nullptr;  // timing: recorded on the launch's stream right before the x-stage kernel
};

// mid (optional): recorded between the x-stage and the epilogue kernel when
// the last stage runs split.  Returns whether it did (then the epilogue ran
// on fork.ep when that is set).
template <typename T>
bool launch_separable_x(const SepParams& sp, const CoupParams& p, bool reduce, hipStream_t s,
                        hipEvent_t mid, const EpFork& fork = EpFork{}) {
  PHX_REQUIRE(p.nm <= kSepNm && p.nb2max <= kSepNb2,
              "separable transform supports nm <= 32 and at most 8 bands per inner point");
  PHX_REQUIRE(!p.fuseYX || (sizeof(T) == 4 && reduce), "separable transform: fused y/x-stage is FP32 only");
  if (sp.nPanels <= 0) return false;
  const int tiles = sp.nPanels * sp.my * sp.mz;
  const int grid = ceil_div(tiles, 8) * 8;
  const int nmb = p.nm <= 12 ? 12 : p.nm <= 16 ? 16 : p.nm <= 24 ? 24 : kSepNm;
  const bool one = p.nCalc == 1;
  constexpr bool bf3t = sizeof(T) == 4;  // split-BF16x3 exists for the FP32 data path
  const bool bf3 = p.bf16x3 != 0;
  constexpr bool bregt = sizeof(T) == 4;  // B in registers for the whole item: FP32, <= 24 x-groups
  const bool breg = bregt && (2 * p.nG0 + 3) / 4 <= 12;
  PHX_REQUIRE(!bf3 || (bf3t && reduce && sp.ncol <= 4 * 6 * 16),
              "split-BF16x3: FP32 data, the linewidth reduction and nm nw <= 384");
  if (reduce && p.screen > 0. && p.screenOf) {
    PHX_REQUIRE(p.xminPart && p.meshToRow, "energy screen: closest-approach buffers missing");
    hipLaunchKernelGGL(k_sep_xmin, dim3(kXminBlocks, sp.nPanels), dim3(256), 0, s, p);
    PHX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sep_xmin_fin, dim3(ceil_div(sp.nPanels, 64)), dim3(64), 0, s, p);
    PHX_HIP(hipGetLastError());
  }
  if (reduce && sp.ncol <= 4 * 6 * 16) {
    const bool wide = sp.ncol > 4 * 3 * 16;
    const size_t ldsXm = sep_xm_lds_bytes(sizeof(T), p.nG0, p.m0, p.nw, p.nm, p.nb2max, nmb <= 16 ? 16 : 8,
                                             p.nCalc);
    PHX_REQUIRE(ldsXm <= 160 * 1024, "separable transform: LDS budget exceeded");
    PHX_REQUIRE(p.qModeGrp && p.E3T, "separable transform: mode-group or eigenvector table missing");
    if (p.keepMask) {
      PHX_REQUIRE(p.m0 <= 64 && p.meshRow, "separable transform: row masks need m0 <= 64");
      hipLaunchKernelGGL(k_sep_screen, dim3(ceil_div(tiles, 256)), dim3(256), 0, s, p);
      PHX_HIP(hipGetLastError());
    }
    if constexpr (sizeof(T) == 4) {
      // the lean epilogue also takes slot masks and a gate (the learned screen's passes)
      const bool ep2ok = !p.epFirst && !p.adaptive && p.nb2max <= 8;
      const bool split = p.G3 && !bf3 && !wide && p.keepMask && (ep2ok || (!p.slotOn && !p.gate)) &&
                         p.nm <= 16 && p.nw <= 16 && (2 * p.nG0 + 3) / 4 <= 12;
      PHX_REQUIRE(split || !p.fuseYX, "separable transform: fused y/x-stage outside the split last stage");
      if (split) {
        // the electron rotation in the x-stage kernel: with the lean epilogue only
        const bool rot = p.epRot && ep2ok;
        const size_t ldsA = sep_xg_lds_bytes(sizeof(T), p.nG0, p.m0, p.nm, p.nw, rot, p.nb2max);
        // whole rows of nm threads per wave, at most 8, fewer if the LDS
        // budget needs it (4 rows per wave always fit what 16 x 16 threads did)
        CoupParams pe = p;
        int rpw = std::min(64 / p.nm, 8);
        constexpr int nWaves = kSepEpThreads / 64;
        while (rpw > 4 && sep_ep_lds_bytes(p.nw, p.nm, p.nb2max, p.nCalc, nWaves * rpw) > 160 * 1024) --rpw;
        pe.epRows = nWaves * rpw;
        const size_t ldsB = sep_ep_lds_bytes(p.nw, p.nm, p.nb2max, p.nCalc, pe.epRows);
        PHX_REQUIRE(ldsA <= 160 * 1024 && ldsB <= 160 * 1024, "separable transform: LDS budget exceeded");
        auto xg = [&](auto kern) {
          PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsA)));
          hipLaunchKernelGGL(kern, dim3(grid), dim3(256), ldsA, s, p);
        };
        if (fork.xgStart) PHX_HIP(hipEventRecord(fork.xgStart, s));
        if (p.fuseYX) {
          if constexpr (sizeof(T) == 4) {
            PHX_REQUIRE(!rot && !p.slotOn && !p.gate, "separable transform: fused y/x-stage without slot masks");
            const size_t ldsYX = sep_yx_lds_bytes(sp.nLines, sp.nG0, p.m0);
            PHX_REQUIRE(ldsYX <= 160 * 1024, "separable transform: LDS budget exceeded");
            PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_sep_yx),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsYX)));
            hipLaunchKernelGGL(k_sep_yx, dim3(ceil_div(sp.ncol, 16), sp.mz, sp.nPanels), dim3(256), ldsYX, s, sp,
                               p);
          }
        } else if (rot) {
          xg(k_sep_xg<T, 3, true>);
        } else {
          xg(k_sep_xg<T, 3, false>);
        }
        PHX_HIP(hipGetLastError());
        if (mid) PHX_HIP(hipEventRecord(mid, s));
        hipStream_t se = s;
        if (fork.ep) {
          PHX_HIP(hipEventRecord(fork.xgDone, s));
          PHX_HIP(hipStreamWaitEvent(fork.ep, fork.xgDone, 0));
          se = fork.ep;
        }
        if (fork.epStart) PHX_HIP(hipEventRecord(fork.epStart, se));
        auto ep = [&](auto kern) {
          PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsB)));
          hipLaunchKernelGGL(kern, dim3(grid), dim3(kSepEpThreads), ldsB, se, pe);
        };
        // partial sums held per thread: calculations x b1-group members of
        // the launch, no more than the launch needs
        const int gm = p.gsizeMax > 0 ? p.gsizeMax : kGMax;
        PHX_REQUIRE(gm <= kGMax, "separable transform: b1 group larger than kGMax");
        if (ep2ok) {
          // lean epilogue: rows per pass capped by its U2 prefetch registers
          const int nb2 = p.nb2max <= 4 ? 4 : 8;
          const int rpw2 = std::min(64 / p.nm, nb2 <= 4 ? 5 : 4);
          const int cm = one ? 1 : p.nCalc <= 2 ? 2 : kCMax;
          const size_t lds2 = sep_ep2_lds_bytes(p.nm, nb2, cm, rpw2);
          PHX_REQUIRE(lds2 <= 160 * 1024, "separable transform: LDS budget exceeded");
          EpParams q{};
          q.keepMask = p.keepMask;
          q.partial = p.partial;
          q.panelKK = p.panelKK;
          q.panelGStart = p.panelGStart;
          q.panelGSize = p.panelGSize;
          q.kkMesh = p.kkMesh;
          q.kkStateOff = p.kkStateOff;
          q.meshRow = p.meshRow;
          q.e1pool = p.e1pool;
          q.e2pool = p.e2pool;
          q.Wpool = p.Wpool;
          q.qModeGrp = p.qModeGrp;
          q.U2pool = p.U2pool;
          q.fermi = p.fermi;
          q.bose = p.bose;
          q.E3T = p.E3T;
          q.G3 = p.G3;
          q.screenOf = p.screenOf;
          q.keptRows = p.keptRows;
          q.keptBands = p.keptBands;
          q.slotOn = p.slotOn;
          q.gate = p.gate;
          q.slotZero = p.slotZero;
          q.invSigma = p.invSigma;
          q.deltaPref = p.deltaPref;
          q.norm = p.norm;
          q.cutoff = p.cutoff;
          q.screen = p.screen;
          q.m0 = p.m0;
          q.m1 = p.m1;
          q.m2 = p.m2;
          q.nPanels = p.nPanels;
          q.nm = p.nm;
          q.nw = p.nw;
          q.nCalc = p.nCalc;
          q.numStates = p.numStates;
          q.rpw = rpw2;
          q.nb2max = p.nb2max;
          auto ep2 = [&](auto kern) {
            PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2)));
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds2, se, q);
          };
          auto ep2f = [&](auto nmx, auto nb2c, auto fz) {
            constexpr int NM = decltype(nmx)::value, NB = decltype(nb2c)::value;
            constexpr bool F = decltype(fz)::value;
            if (gm <= 4) {
              if (one) ep2(k_sep_ep2<T, 1, NM, 4, NB, F>);
              else if (p.nCalc <= 2) ep2(k_sep_ep2<T, 2, NM, 4, NB, F>);
              else ep2(k_sep_ep2<T, kCMax, NM, 4, NB, F>);
            } else {
              if (one) ep2(k_sep_ep2<T, 1, NM, kGMax, NB, F>);
              else if (p.nCalc <= 2) ep2(k_sep_ep2<T, 2, NM, kGMax, NB, F>);
              else ep2(k_sep_ep2<T, kCMax, NM, kGMax, NB, F>);
            }
          };
          auto ep2c = [&](auto nmx, auto nb2c) {
            if (rot) ep2f(nmx, nb2c, std::true_type{});
            else ep2f(nmx, nb2c, std::false_type{});
          };
          using I12 = std::integral_constant<int, 12>;
          using I16 = std::integral_constant<int, 16>;
          using B4 = std::integral_constant<int, 4>;
          using B8 = std::integral_constant<int, 8>;
          if (p.nm <= 12) nb2 <= 4 ? ep2c(I12{}, B4{}) : ep2c(I12{}, B8{});
          else nb2 <= 4 ? ep2c(I16{}, B4{}) : ep2c(I16{}, B8{});
          PHX_HIP(hipGetLastError());
          return true;
        }
        auto epc = [&](auto nmx) {
          constexpr int NM = decltype(nmx)::value;
          if (gm <= 4) {
            if (one) ep(k_sep_ep<T, 1, NM, 4>);
            else if (p.nCalc <= 2) ep(k_sep_ep<T, 2, NM, 4>);
            else ep(k_sep_ep<T, kCMax, NM, 4>);
          } else {
            if (one) ep(k_sep_ep<T, 1, NM, kGMax>);
            else if (p.nCalc <= 2) ep(k_sep_ep<T, 2, NM, kGMax>);
            else ep(k_sep_ep<T, kCMax, NM, kGMax>);
          }
        };
        if (p.nm <= 12) epc(std::integral_constant<int, 12>{});
        else epc(std::integral_constant<int, 16>{});
        PHX_HIP(hipGetLastError());
        return true;
      }
    }
    auto go = [&](auto kern) {
      PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsXm)));
      hipLaunchKernelGGL(kern, dim3(grid), dim3(256), ldsXm, s, p);
    };
#define PHX_XM(NMX)                                                                 \
    if (nmb == NMX) {                                                               \
      if (wide) {                                                                   \
        if (bf3 && one) go(k_sep_xm<T, true, NMX, 1, bf3t, 6>);                      \
        else if (bf3) go(k_sep_xm<T, true, NMX, kCMax, bf3t, 6>);                    \
        else if (one) go(k_sep_xm<T, true, NMX, 1, false, 6>);                       \
        else go(k_sep_xm<T, true, NMX, kCMax, false, 6>);                            \
      } else {                                                                      \
        if (bf3 && one) go(k_sep_xm<T, true, NMX, 1, bf3t>);                         \
        else if (bf3) go(k_sep_xm<T, true, NMX, kCMax, bf3t>);                       \
        else if (breg && one) go(k_sep_xm<T, true, NMX, 1, false, 3, bregt>);        \
        else if (breg) go(k_sep_xm<T, true, NMX, kCMax, false, 3, bregt>);           \
        else if (one) go(k_sep_xm<T, true, NMX, 1>);                                 \
        else go(k_sep_xm<T, true, NMX, kCMax>);                                      \
      }                                                                             \
    }
    PHX_XM(12) PHX_XM(16) PHX_XM(24) PHX_XM(32)
#undef PHX_XM
    PHX_HIP(hipGetLastError());
    return false;
  }
  PHX_REQUIRE(!p.slotOn && !p.gate, "separable transform: slot masks need nm nw <= 384");
  const size_t ldsX = std::max(sep_x_lds_bytes(sizeof(T), p.nG0, sp.ncol, p.m0, p.nw, p.nm, p.nb2max),
                               size_t(kGMax) * kCMax * 4 * sizeof(double));
  PHX_REQUIRE(ldsX <= 160 * 1024, "separable transform: LDS budget exceeded");
  auto go = [&](auto kern) {
    PHX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsX)));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), ldsX, s, p);
  };
  const int nmo = p.nm <= 12 ? 12 : p.nm <= 18 ? 18 : kSepNm;
  if (reduce) {
    if (nmo == 12) one ? go(k_sep_x<T, true, 12, 1>) : go(k_sep_x<T, true, 12, kCMax>);
    else if (nmo == 18) one ? go(k_sep_x<T, true, 18, 1>) : go(k_sep_x<T, true, 18, kCMax>);
    else one ? go(k_sep_x<T, true, kSepNm, 1>) : go(k_sep_x<T, true, kSepNm, kCMax>);
  } else {
    if (nmo == 12) go(k_sep_x<T, false, 12, 1>);
    else if (nmo == 18) go(k_sep_x<T, false, 18, 1>);
    else go(k_sep_x<T, false, kSepNm, 1>);
  }
  PHX_HIP(hipGetLastError());
  return false;
}

template <typename T>
void launch_separable(const SepParams& sp, const CoupParams& p, bool reduce, int maxLineE,
                      int maxGrpE, hipStream_t s) {
  launch_separable_zy<T>(sp, maxLineE, maxGrpE, s);
  launch_separable_x<T>(sp, p, reduce, s, nullptr);
}

}  // namespace phx
