// Parameter blocks and launchers of the cacheElPh and collect kernels
// (elph.hip), shared by the kernels and the host side (api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace phx {

constexpr int kGMax = 8;       // largest degenerate b1 group handled in one panel
constexpr int kCMax = 4;       // largest number of (T, mu) calculations
constexpr int kCacheNKK = 32;  // outer points per GEMM cacheElPh launch
constexpr int kLinesPerBatch = 16;   // mesh lines per line-form cacheElPh launch
constexpr int kLinePointsMax = 1024; // outer points per line-form batch

// Line counts k_cache_line_a is instantiated for; a batch's count is rounded
// up to the next one (the extra lines have zero phases and are never read).
inline int line_count_padded(int n) {
  return n <= 4 ? n : n <= 6 ? 6 : n <= 8 ? 8 : n <= 12 ? 12 : 16;
}

// g1(k1; Rp,nu,w1,w2) = sum_Re e^{ik1.Re}/de gW(Re,Rp,nu,w1,w2), rotated by U1
// and scaled by s(k1, Rp) into one real panel per (k1, b1).
struct CacheParams {
  int nRe, nRp, nm, nw, nkk;
  const void* gW;         // (nRe, nRp, nm, nw, nw) complex, element type T
  const double2* phE;     // (nkk, nRe): e^{ik1.Re}/de (GEMM form)
  const double2* phP;     // (nkk, nRp): e^{-ik1.Rp}/dp or 1/dp
  const double2* U1pool;  // per kk: (nw, nb1) column-major
  const int* U1off;       // complex offset per kk
  const int* nb1;         // per kk
  const int* panelBase;   // first panel of kk
  void* B;                // panels
  size_t panelStride;     // elements
  int ldb;                // elements per panel row
  // outer points on one mesh line (k_cache_line_a / k_cache_line_b)
  size_t N;               // gW columns: nRp * nm * nw^2
  int nGe;                // R_e groups: distinct n0(R_e) mod m0
  const int* reOrder;     // nRe: native index of the i-th R_e in group order
  const int* reGroupEnd;  // nRe: group index if the i-th R_e ends its group, else -1
  // the same order with every group padded to a multiple of 4 entries (k_cache_line_am)
  int nRePad;
  const int2* rePad;      // nRePad: (gW row, position in group order) of each entry, (-1, -1) = padding
  const int* reBlkEnd;    // nRePad / 4: group index if the 4-entry block ends its group, else -1
  int nLinesB;            // lines of the batch (<= kLinesPerBatch)
  int nLinesA;            // lines stage a computes: line_count_padded(nLinesB), rows of phLine and Tg
  const int* lineKK;      // nLinesB + 1: first outer point of each line (points grouped by line)
  const void* phLine;     // (nRe, nLinesA) complex T: e^{2 pi i (j1 n1/m1 + j2 n2/m2)} / de
  void* Tg;               // (nLinesA, nGe, N) complex T: partial sums of the lines
  const double2* wLine;   // (nkk, nGe): e^{2 pi i j0(kk) n0(g) / m0}
};

// Deterministic sum of the per-tile partial linewidths of one batch.
struct CollectParams {
  int nkk, nMtiles, nCalc, numStates;
  int nb1Max;  // largest nb1 of the batch
  const int* kkNb1; const int* kkPanelBase; const int* kkStateOff;
  const int* panelGStart; const int* panelGSize;
  const double* partial;
  double* linewidth;  // (nCalc, numStates), accumulated into
};

template <typename T, typename OutT> void launch_cache(const CacheParams& p, hipStream_t s);
template <typename T, typename OutT> void launch_cache_line(const CacheParams& p, hipStream_t s);
void launch_collect(const CollectParams& p, hipStream_t s);

}  // namespace phx
