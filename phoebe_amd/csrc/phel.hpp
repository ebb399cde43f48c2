// Reduction half of the phonon-linewidth builder (phel.hip): parameters and launcher.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace phx {

struct PhelParams {
  int nq, nb1, nb2, nm, nCalc;
  double* g2;             // per q3 a block of g2Stride, (b1, b2, nu) column-major inside
  size_t g2Stride;
  const double* e1;       // nb1 (the k1 bands)
  const double* E2;       // band energies of the k2 points, nb2 per point
  const int* e2off;       // per q3: offset of its k2 point's energies in E2
  const double* W3;       // nq x nm
  const double* F1;       // nCalc x nb1: f (1 - f) of the k1 states
  const double* invT;     // nCalc
  double norm, cutoff, invSigma, deltaPref;
  int adaptive;           // width from v1 - v2 (AdaptiveGaussianDeltaFunction)
  const double* v1;       // nb1 x 3
  const double* V2;       // nq x nb2 x 3
  double qT[9], minWidth;
  double* lw;             // nCalc x nq x nm
};

void launch_phel_reduce(const PhelParams& p, hipStream_t s);
void launch_phel_flag_failures(int n, const int* info, int* flag, hipStream_t s);

}  // namespace phx
