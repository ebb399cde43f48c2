// Learned q-block screen: MLP forward / backward / Adam on
// v_mfma_f32_16x16x32_bf16, block features, skip decisions and the audit.
// No counterpart in the reference, which evaluates every (k1, k2) pair
// (el_scattering.cpp:161-386).
#include <algorithm>
#include <cmath>

#include "elph_params.hpp"
#include "phx_common.hpp"
#include "screen.hpp"

namespace phx {

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ GEMM
// Workgroup tile 64 x 64: wave w owns rows 16w..16w+15 and four 16-column
// tiles.  gfx950's v_mfma_f32_16x16x32_bf16: A lane l holds row l%16, B lane l
// column l%16, each with 8 K values (here k = 8(l/16) + j, the same slots in
// A and B, which is all the product needs); D lane l holds D[4(l/16)+i][l%16].
__global__ void __launch_bounds__(256) k_mlp_gemm(MlpGemm g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int m0 = blockIdx.x * 64 + wave * 16, n0 = blockIdx.y * 64;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int m = m0 + lr;
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    s16x8 a;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + 8 * lg + j;
      float v = 0.f;
      if (m < g.M && k < g.K) v = g.ta ? g.A[size_t(k) * g.lda + m] : g.A[size_t(m) * g.lda + k];
      a[j] = bf16_bits(v);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int n = n0 + 16 * t + lr;
      s16x8 b;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + 8 * lg + j;
        float v = 0.f;
        if (n < g.N && k < g.K) v = g.tb ? g.B[size_t(n) * g.ldb + k] : g.B[size_t(k) * g.ldb + n];
        b[j] = bf16_bits(v);
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int mm = m0 + 4 * lg + i, n = n0 + 16 * t + lr;
      if (mm < g.M && n < g.N) {
        float v = acc[t][i];
        if (g.bias) v += g.bias[n];
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.mask && !(g.mask[size_t(mm) * g.ldm + n] > 0.f)) v = 0.f;
        g.C[size_t(mm) * g.ldc + n] = v;
      }
    }
}

void launch_mlp_gemm(const MlpGemm& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  hipLaunchKernelGGL(k_mlp_gemm, dim3(ceil_div(g.M, 64), ceil_div(g.N, 64)), dim3(256), 0, s, g);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ small kernels
// out[c] = sum_m Z[m][c] (one workgroup per column, fixed-order tree)
__global__ void __launch_bounds__(256) k_colsum(const float* Z, int n, int ld, float* out) {
  __shared__ float red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  float s = 0.f;
  for (int m = t; m < n; m += 256) s += Z[size_t(m) * ld + c];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[c] = red[0];
}

// dy = 2 (y - t) / n, loss = mean (y - t)^2 (one workgroup)
__global__ void __launch_bounds__(256) k_mse(const float* y, const float* tg, int n, float* dy, float* loss) {
  __shared__ float red[256];
  const int t = threadIdx.x;
  float s = 0.f;
  for (int m = t; m < n; m += 256) {
    const float d = y[m] - tg[m];
    dy[m] = 2.f * d / float(n);
    s += d * d;
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) loss[0] = red[0] / float(n);
}

__global__ void k_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// ------------------------------------------------------------------ MLP
size_t mlp_scratch_floats(int batch) { return size_t(batch) * (4 * kMlpHid + 2) + 64; }

// scratch: H1 (n x 64), H2 (n x 64), dZ2 (n x 64), dZ1 (n x 64), dy (n)
static MlpGemm gemm(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb,
                    int tb, float* C, int ldc) {
  MlpGemm g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.ta = ta;
  g.B = B; g.ldb = ldb; g.tb = tb;
  g.C = C; g.ldc = ldc;
  return g;
}

void mlp_forward(const float* P, const float* X, int n, float* scratch, float* y, hipStream_t s) {
  float* H1 = scratch;
  float* H2 = H1 + size_t(n) * kMlpHid;
  MlpGemm g = gemm(n, kMlpHid, kMlpIn, X, kMlpIn, 0, P + kMlpOffW1, kMlpHid, 0, H1, kMlpHid);
  g.bias = P + kMlpOffB1;
  g.relu = 1;
  launch_mlp_gemm(g, s);
  g = gemm(n, kMlpHid, kMlpHid, H1, kMlpHid, 0, P + kMlpOffW2, kMlpHid, 0, H2, kMlpHid);
  g.bias = P + kMlpOffB2;
  g.relu = 1;
  launch_mlp_gemm(g, s);
  g = gemm(n, 1, kMlpHid, H2, kMlpHid, 0, P + kMlpOffW3, 1, 0, y, 1);
  g.bias = P + kMlpOffB3;
  launch_mlp_gemm(g, s);
}

void mlp_backward(const float* P, const float* X, const float* t, int n, float* scratch,
                  float* G, float* loss, hipStream_t s) {
  float* H1 = scratch;
  float* H2 = H1 + size_t(n) * kMlpHid;
  float* dZ2 = H2 + size_t(n) * kMlpHid;
  float* dZ1 = dZ2 + size_t(n) * kMlpHid;
  float* dy = dZ1 + size_t(n) * kMlpHid;
  float* y = dy + n;
  mlp_forward(P, X, n, scratch, y, s);
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, s, y, t, n, dy, loss);
  PHX_HIP(hipGetLastError());
  // layer 3: dW3 = H2^T dy, db3 = sum dy, dZ2 = (dy W3^T) * (H2 > 0)
  launch_mlp_gemm(gemm(kMlpHid, 1, n, H2, kMlpHid, 1, dy, 1, 0, G + kMlpOffW3, 1), s);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, s, dy, n, 1, G + kMlpOffB3);
  MlpGemm g = gemm(n, kMlpHid, 1, dy, 1, 0, P + kMlpOffW3, 1, 1, dZ2, kMlpHid);
  g.mask = H2;
  g.ldm = kMlpHid;
  launch_mlp_gemm(g, s);
  // layer 2: dW2 = H1^T dZ2, db2 = colsum dZ2, dZ1 = (dZ2 W2^T) * (H1 > 0)
  launch_mlp_gemm(gemm(kMlpHid, kMlpHid, n, H1, kMlpHid, 1, dZ2, kMlpHid, 0, G + kMlpOffW2, kMlpHid), s);
  hipLaunchKernelGGL(k_colsum, dim3(kMlpHid), dim3(256), 0, s, dZ2, n, kMlpHid, G + kMlpOffB2);
  g = gemm(n, kMlpHid, kMlpHid, dZ2, kMlpHid, 0, P + kMlpOffW2, kMlpHid, 1, dZ1, kMlpHid);
  g.mask = H1;
  g.ldm = kMlpHid;
  launch_mlp_gemm(g, s);
  // layer 1: dW1 = X^T dZ1, db1 = colsum dZ1
  launch_mlp_gemm(gemm(kMlpIn, kMlpHid, n, X, kMlpIn, 1, dZ1, kMlpHid, 0, G + kMlpOffW1, kMlpHid), s);
  hipLaunchKernelGGL(k_colsum, dim3(kMlpHid), dim3(256), 0, s, dZ1, n, kMlpHid, G + kMlpOffB1);
  PHX_HIP(hipGetLastError());
}

void mlp_adam(float* P, const float* G, float* m1, float* m2, int step, float lr, hipStream_t s) {
  const float c1 = 1.f - std::pow(0.9f, float(step)), c2 = 1.f - std::pow(0.999f, float(step));
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(kMlpParams, 256)), dim3(256), 0, s, P, G, m1, m2, kMlpParams,
                     lr, c1, c2);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ features
// Distance of 0 from the interval [lo, hi].
__device__ inline double interval_gap(double lo, double hi) {
  return lo > 0. ? lo : (hi < 0. ? -hi : 0.);
}

__global__ void __launch_bounds__(256) k_block_features(FeatParams p) {
  const int nLinesYZ = p.m1 * p.m2;
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= p.nPanels * nLinesYZ) return;
  const int panel = slot / nLinesYZ, yz = slot % nLinesYZ;
  const int jy = yz % p.m1, jz = yz / p.m1;
  const int kk = p.panelKK[panel];
  const int kmesh = p.kkMesh[kk];
  int j0, j1, j2;
  mesh_coords(kmesh, p.m0, p.m1, j0, j1, j2);
  const double* e1g = p.epool + p.kkStateOff[kk] + p.panelGStart[panel];
  const int gsz = p.panelGSize[panel];
  const double e1 = p.epool[p.kkStateOff[kk] + p.panelB1[panel]];
  double dmin = 1e300, e2min = 1e300, e2max = -1e300, wmin = 1e300, wmax = -1e300, wsum = 0.;
  int n3 = 0, nscr = 0, nval = 0;
  for (int jx = 0; jx < p.m0; ++jx) {
    const int mesh = (jz * p.m1 + jy) * p.m0 + jx;
    const int row = p.meshToRow[mesh];
    if (row < p.rowLo || row >= p.rowHi) continue;
    ++nval;
    const int iq = mesh_diff_index(mesh, kmesh, p.m0, p.m1, p.m2);
    const double wlo = p.qWlo[iq], whi = p.qWhi[iq];
    const double2 er = p.rowE2range[row];
    e2min = fmin(e2min, er.x);
    e2max = fmax(e2max, er.y);
    if (!(whi >= wlo)) continue;  // no mode above the cutoff
    wmin = fmin(wmin, wlo);
    wmax = fmax(wmax, whi);
    double d = 1e300;
    for (int gi = 0; gi < gsz; ++gi) {
      d = fmin(d, interval_gap(e1g[gi] - er.y + wlo, e1g[gi] - er.x + whi));
      d = fmin(d, interval_gap(e1g[gi] - er.y - whi, e1g[gi] - er.x - wlo));
    }
    dmin = fmin(dmin, d);
    n3 += d < 3. * p.sigma;
    if (d < p.screen * p.sigma) {
      ++nscr;
      // the row's largest Gaussian weight: its interval gap to resonance
      const double x = d / p.sigma;
      wsum += exp(-x * x);
    }
  }
  if (nval == 0) e2min = e2max = p.mu;
  if (!(wmax >= wmin)) wmin = wmax = 0.;
  float* f = p.feat + size_t(slot) * kMlpIn;
  f[0] = float(j0) / p.m0;
  f[1] = float(j1) / p.m1;
  f[2] = float(j2) / p.m2;
  f[3] = float(((jy - j1) % p.m1 + p.m1) % p.m1) / p.m1;
  f[4] = float(((jz - j2) % p.m2 + p.m2) % p.m2) / p.m2;
  f[5] = float((e1 - p.mu) / p.escale);
  f[6] = float((e2min - p.mu) / p.escale);
  f[7] = float((e2max - p.mu) / p.escale);
  f[8] = float(wmin / p.wscale);
  f[9] = float(wmax / p.wscale);
  f[10] = float(fmin(dmin / p.sigma, 10.) / 10.);
  f[11] = float(log1p(double(n3)) / log1p(double(p.m0)));
  f[12] = float(nscr) / p.m0;
  f[13] = float(nval) / p.m0;
  f[14] = float(gsz) / kGMax;
  // ln of the summed row weights: the baseline the network's prediction is
  // added to (it learns ln contribution - baseline), stored scaled
  f[15] = float((wsum > 0. ? log(wsum) : -kScreenBaseFloor) / kScreenBaseScale);
  // rows inside the energy screen (low 16 bits) and within 3 sigma of resonance (high bits)
  p.nscreen[slot] = nscr | (n3 << 16);
}

void launch_block_features(const FeatParams& p, hipStream_t s) {
  const int n = p.nPanels * p.m1 * p.m2;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_block_features, dim3(ceil_div(n, 256)), dim3(256), 0, s, p);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ decisions
// The block's baseline ln contribution (feature 15, scaled back).
__device__ inline double screen_base(const float* feat, size_t slot) {
  return double(feat[slot * kMlpIn + 15]) * kScreenBaseScale;
}

__global__ void __launch_bounds__(256) k_screen_decide(DecideParams p) {
  __shared__ double red[256];
  __shared__ long long cnt[3];
  const int panel = blockIdx.x, t = threadIdx.x;
  const size_t s0 = size_t(panel) * p.nLinesYZ;
  double sum = 0.;
  for (int yz = t; yz < p.nLinesYZ; yz += 256)
    if ((p.nscreen[s0 + yz] & 0xffff) > 0) sum += exp(double(p.pred[s0 + yz]) + screen_base(p.feat, s0 + yz));
  red[t] = sum;
  if (t < 3) cnt[t] = 0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double thrAbs = p.thr * red[0];
  if (t == 0) p.thrAbs[panel] = thrAbs;
  long long nin = 0, nskip = 0, naud = 0;
  for (int yz = t; yz < p.nLinesYZ; yz += 256) {
    const size_t slot = s0 + yz;
    unsigned char c = 0, a = 0, k = 0;
    const int nscr = p.nscreen[slot] & 0xffff, nnear = p.nscreen[slot] >> 16;
    if (nscr > 0) {
      ++nin;
      // blocks with a row within 3 sigma of resonance are always exact
      if (nnear > 0 || exp(double(p.pred[slot]) + screen_base(p.feat, slot) + p.margin) >= thrAbs) {
        c = 1;
      } else {
        ++nskip;
        if (screen_draw(p.seed, p.batchId, p.panelBase + panel, p.nLinesYZ, yz) < p.auditFrac) {
          a = 1;
          ++naud;
        } else {
          k = 1;
        }
      }
    }
    p.computeMask[slot] = c;
    p.auditMask[slot] = a;
    p.skipMask[slot] = k;
  }
  atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[0]), (unsigned long long)nin);
  atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[1]), (unsigned long long)nskip);
  atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[2]), (unsigned long long)naud);
  __syncthreads();
  if (t < 3) atomicAdd(reinterpret_cast<unsigned long long*>(p.counts + t), (unsigned long long)cnt[t]);
}

void launch_screen_decide(const DecideParams& p, hipStream_t s) {
  if (p.nPanels <= 0) return;
  hipLaunchKernelGGL(k_screen_decide, dim3(p.nPanels), dim3(256), 0, s, p);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ audit
// One workgroup per panel: exact contributions of the audited blocks (calc 0,
// all members of b1's group) against the panel's threshold; their partials are
// exact and replace the zeros of the primary pass.  Per-panel statistics go to
// panelStats and are folded in fixed order by k_audit_fold.
__global__ void __launch_bounds__(256) k_screen_audit(AuditParams p, double* panelStats) {
  __shared__ double rmax[256], rsum[256];
  const int panel = blockIdx.x, t = threadIdx.x;
  const size_t s0 = size_t(panel) * p.nLinesYZ;
  const double thr = p.thrAbs[panel];
  double mx = 0., sm = 0.;
  for (int yz = t; yz < p.nLinesYZ; yz += 256) {
    const size_t slot = s0 + yz;
    if (!p.auditMask[slot]) continue;
    const double* src = p.partialAudit + slot * kGMax * kCMax;
    double c = 0.;
    for (int gi = 0; gi < kGMax; ++gi) c += src[gi * kCMax];
    const double r = thr > 0. ? c / thr : (c > 0. ? 1e300 : 0.);
    mx = fmax(mx, r);
    sm += r;
    if (r > p.failFactor) *p.fail = 1;
    double* dst = p.partial + slot * kGMax * kCMax;
    for (int i = 0; i < kGMax * kCMax; ++i) dst[i] = src[i];
  }
  rmax[t] = mx;
  rsum[t] = sm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      rmax[t] = fmax(rmax[t], rmax[t + w]);
      rsum[t] += rsum[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    panelStats[2 * panel] = rmax[0];
    panelStats[2 * panel + 1] = rsum[0];
  }
}

__global__ void k_audit_fold(int nPanels, const double* panelStats, double* stats) {
  if (threadIdx.x != 0) return;
  double mx = stats[0], sm = stats[1];
  for (int i = 0; i < nPanels; ++i) {
    mx = fmax(mx, panelStats[2 * i]);
    sm += panelStats[2 * i + 1];
  }
  stats[0] = mx;
  stats[1] = sm;
}

void launch_screen_audit(const AuditParams& p, double* panelStats, hipStream_t s) {
  if (p.nPanels <= 0) return;
  hipLaunchKernelGGL(k_screen_audit, dim3(p.nPanels), dim3(256), 0, s, p, panelStats);
  PHX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_audit_fold, dim3(1), dim3(64), 0, s, p.nPanels, panelStats, p.stats);
  PHX_HIP(hipGetLastError());
}

__global__ void k_fail_fold(int* fail, long long* fallbacks) {
  if (threadIdx.x == 0) {
    fallbacks[0] += *fail;
    *fail = 0;
  }
}

void launch_fail_fold(int* fail, long long* fallbacks, hipStream_t s) {
  hipLaunchKernelGGL(k_fail_fold, dim3(1), dim3(64), 0, s, fail, fallbacks);
  PHX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ training targets
// X[i] = feat[slot_i], t[i] = ln max(c_i, 1e-40) with c the block's calc-0
// contribution summed over b1's degenerate group.
__global__ void k_gather_targets(int nSel, const int* slots, const double* partial, const float* feat,
                                 float* X, float* t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nSel) return;
  const size_t slot = size_t(slots[i]);
  const double* src = partial + slot * kGMax * kCMax;
  double c = 0.;
  for (int gi = 0; gi < kGMax; ++gi) c += src[gi * kCMax];
  // residual target: ln contribution minus the block's Gaussian baseline
  t[i] = float(log(fmax(c, 1e-40)) - double(feat[slot * kMlpIn + 15]) * kScreenBaseScale);
  for (int f = 0; f < kMlpIn; ++f) X[size_t(i) * kMlpIn + f] = feat[slot * kMlpIn + f];
}

void launch_gather_targets(int nSel, const int* slots, const double* partial, const float* feat,
                           float* X, float* t, hipStream_t s) {
  if (nSel <= 0) return;
  hipLaunchKernelGGL(k_gather_targets, dim3(ceil_div(nSel, 256)), dim3(256), 0, s, nSel, slots, partial,
                     feat, X, t);
  PHX_HIP(hipGetLastError());
}

}  // namespace phx
