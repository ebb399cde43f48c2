// Learned q-block screen of the separable transform: parameter blocks and
// launchers (screen.hip), shared with the host side (api.hip).
//
// A q-block is one work item of the last transform stage: one panel (k1, b1)
// and one (jy, jz) line of inner points k2 (all jx).  A 3-layer MLP
// (kMlpIn -> 64 -> 64 -> 1, BF16 operands, FP32 accumulation) predicts the
// natural log of the block's contribution to the linewidths of b1's
// degenerate group from features of k1, the q-line, the band energies and the
// phonon frequencies.  Blocks predicted below a fraction of the panel's
// predicted total are skipped; a random fraction of the skipped blocks is
// audited through the exact kernel, and a failed audit recomputes every
// skipped block of the batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace phx {

constexpr int kMlpIn = 16;   // features per block
constexpr int kMlpHid = 64;  // hidden width
// parameter layout (FP32 master copy): W1 (16 x 64), b1 (64), W2 (64 x 64), b2 (64), W3 (64 x 1), b3 (1)
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHid;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHid;
constexpr int kMlpParams = kMlpOffB3 + 1;
// Feature 15 of a block is ln(sum over its rows inside the energy screen of
// e^{-x^2}), x the row's closest approach to resonance in units of sigma,
// divided by kScreenBaseScale; -kScreenBaseFloor when no row is inside.  The
// network predicts ln(contribution) minus this baseline.
constexpr double kScreenBaseScale = 32.;
constexpr double kScreenBaseFloor = 64.;

// C(M x N) = opA(M x K) opB(K x N) with operands rounded to BF16 (round to
// nearest even) and FP32 accumulation, then C += bias(n), relu, or
// C *= (mask(m, n) > 0).  ta: A is stored K x M (element (m, k) at A[k*lda + m]);
// tb: B is stored N x K (element (k, n) at B[n*ldb + k]).
struct MlpGemm {
  int M, N, K;
  const float* A; int lda; int ta;
  const float* B; int ldb; int tb;
  float* C; int ldc;
  const float* bias;
  int relu;
  const float* mask; int ldm;
};
void launch_mlp_gemm(const MlpGemm& g, hipStream_t s);

// Forward / backward / Adam over one minibatch (device pointers; scratch holds
// the activations and gradients: see mlp_scratch_floats).
size_t mlp_scratch_floats(int batch);
void mlp_forward(const float* params, const float* X, int n, float* scratch, float* y, hipStream_t s);
// loss = mean (y - t)^2; writes grads (kMlpParams) and loss (1 float)
void mlp_backward(const float* params, const float* X, const float* t, int n, float* scratch,
                  float* grads, float* loss, hipStream_t s);
void mlp_adam(float* params, const float* grads, float* m1, float* m2, int step, float lr,
              hipStream_t s);

// Per-block features and the count of rows inside the energy screen.
struct FeatParams {
  int nPanels, m0, m1, m2;
  const int* panelKK; const int* panelB1; const int* panelGStart; const int* panelGSize;
  const int* kkMesh; const int* kkStateOff; const double* epool;
  const int* meshToRow; int rowLo, rowHi;
  const double2* rowE2range; const double* qWlo; const double* qWhi;
  double mu, escale, wscale, sigma, screen;
  float* feat;   // (slots, kMlpIn), slot = panel * m1 * m2 + yz
  int* nscreen;  // (slots): rows inside the energy screen | rows within 3 sigma << 16
};
void launch_block_features(const FeatParams& p, hipStream_t s);

// Skip decisions from the predicted log contributions (one workgroup per panel).
// A block is computed when a row lies within 3 sigma of resonance, or when
// exp(pred + margin) >= thr * (sum of exp(pred) over the panel's blocks).
struct DecideParams {
  int nPanels, nLinesYZ;
  const float* feat;       // block features; feature 15 is the block's baseline / kScreenBaseScale
  int panelBase;           // batch-wide index of panel 0 (hash key of the audit draw)
  const float* pred; const int* nscreen;
  double thr, auditFrac; uint64_t seed, batchId;
  double margin;           // added to the predicted ln contribution before the comparison
  unsigned char* computeMask; unsigned char* auditMask; unsigned char* skipMask;
  double* thrAbs;          // per panel: thr * predicted panel total
  long long* counts;       // [blocks with rows in the screen, learned skips, audits], accumulated
};
void launch_screen_decide(const DecideParams& p, hipStream_t s);

// Audit: exact contributions of the audited blocks against the threshold;
// copies their partials into the main buffer and raises *fail on a miss.
struct AuditParams {
  int nPanels, nLinesYZ;
  const unsigned char* auditMask; const double* thrAbs;
  const double* partialAudit; double* partial;  // (slots, kGMax, kCMax)
  double failFactor;
  int* fail;               // set to 1 when an audited block exceeds failFactor * thrAbs
  double* stats;           // [max audited c / thrAbs, sum of audited c / thrAbs], accumulated
};
// panelStats: nPanels x 2 doubles of scratch
void launch_screen_audit(const AuditParams& p, double* panelStats, hipStream_t s);

// Counts a raised fail flag into fallbacks[0] and clears it.
void launch_fail_fold(int* fail, long long* fallbacks, hipStream_t s);

// Uniform [0, 1) draw of block (panel, yz) of batch batchId (audit and training samples).
__host__ __device__ inline uint64_t screen_mix(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}
__host__ __device__ inline double screen_draw(uint64_t seed, uint64_t batchId, int panel, int nLinesYZ,
                                              int yz) {
  const uint64_t key = batchId * 0x9E3779B97F4A7C15ull + uint64_t(panel) * uint64_t(nLinesYZ) + uint64_t(yz);
  return double(screen_mix(seed ^ screen_mix(key)) >> 11) * (1.0 / 9007199254740992.0);
}

// Training targets: log contribution of selected blocks, gathered in slot order.
void launch_gather_targets(int nSel, const int* slots, const double* partial, const float* feat,
                           float* X, float* t, hipStream_t s);

}  // namespace phx
