// Froehlich (polar) term of the el-ph coupling: device parameters and launcher
// of polarCorrectionPart1 (polar.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/phx_elph.h"
#include "phx_common.hpp"

namespace phx {

constexpr int kPolarMaxModes = 36;   // 3 x 12 atoms
constexpr double kPolarGMax = 14.;   // interaction_elph.cpp:163

struct PolarDev {
  int natoms = 0;
  int coarse[3] = {0, 0, 0};
  double recip[9] = {}, eps[9] = {};
  double volume = 0.;
  const double* pos = nullptr;   // natoms x 3 (device)
  const double* born = nullptr;  // natoms x 3 x 3 (device)
};

// The parameters of a handle, with their device copies.
struct PolarHost {
  bool on = false;
  PolarDev d;
  DevBuf<double> dPos, dBorn;
  void set(const phx_polar* p, hipStream_t s);  // null: off
};

// x(q, nu) for nq wavevectors q3 (device, nq x 3) with eigenvectors e3
// (device, nq x (nm x nm) column-major) into out (device, nq x nm).
void launch_polar_part1(const PolarDev& d, int nq, int nm, const double* q3, const double2* e3, double2* out,
                        hipStream_t s);

}  // namespace phx
