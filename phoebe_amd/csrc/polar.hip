// Froehlich long-range part of the el-ph coupling (the reference's
// polarCorrectionPart1Static, interaction_elph.cpp:152-203; Verdi and
// Giustino, PRL 115, 176401, Eq. 4) for a batch of wavevectors.
#include <vector>

#include "polar.hpp"

namespace phx {

// One workgroup per wavevector q.  For each atom a, the threads sum over
// their share of the reciprocal vectors G of the box |m_i| <= coarse[i]:
//   A(3a + pol) = sum_G pref e^{-qe/4} / qe e^{-i (q+G).tau_a} ((q+G).Z(a, :, pol)),
// qe = (q+G).eps.(q+G), over 0 < qe < 4 gMax; then
//   x(nu) = i sum_k A(k) e3(k, nu).
// The partial sums are reduced in a fixed order (wave shuffles, then waves).
__global__ void __launch_bounds__(256) k_polar_part1(PolarDev d, int nm, const double* q3, const double2* e3,
                                                     double2* out) {
  const int iq = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  __shared__ double2 red[4][3];
  __shared__ double2 A[kPolarMaxModes];
  const double q0 = q3[3 * iq], q1 = q3[3 * iq + 1], q2 = q3[3 * iq + 2];
  const int n1 = 2 * d.coarse[1] + 1, n2 = 2 * d.coarse[2] + 1;
  const int nG = (2 * d.coarse[0] + 1) * n1 * n2;
  // chargeSquare * fourPi / volume with chargeSquare = 2 and fourPi = 2 pi
  // (constants.h:12), interaction_elph.cpp:164-166
  const double pref = 2.0 * (2.0 * 3.14159265358979323846) / d.volume;
  for (int a = 0; a < d.natoms; ++a) {
    double2 acc[3] = {make_double2(0., 0.), make_double2(0., 0.), make_double2(0., 0.)};
    const double tx = d.pos[3 * a], ty = d.pos[3 * a + 1], tz = d.pos[3 * a + 2];
    const double* Z = d.born + 9 * a;
    for (int g = t; g < nG; g += 256) {
      const int m0 = g / (n1 * n2) - d.coarse[0];
      const int r = g % (n1 * n2);
      const int m1 = r / n2 - d.coarse[1], m2 = r % n2 - d.coarse[2];
      double G[3];
      for (int i = 0; i < 3; ++i) {
        G[i] = d.recip[i] * m0 + d.recip[i + 3] * m1 + d.recip[i + 6] * m2;
      }
      G[0] += q0;
      G[1] += q1;
      G[2] += q2;
      double qe = 0.;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) qe += G[i] * d.eps[3 * i + j] * G[j];
      if (!(qe > 0. && qe / 4. < kPolarGMax)) continue;
      const double f2 = pref * exp(-qe / 4.) / qe;
      const double arg = -(G[0] * tx + G[1] * ty + G[2] * tz);
      double sn, cs;
      sincos(arg, &sn, &cs);
      for (int pol = 0; pol < 3; ++pol) {
        const double gz = G[0] * Z[pol] + G[1] * Z[3 + pol] + G[2] * Z[6 + pol];
        acc[pol].x += f2 * gz * cs;
        acc[pol].y += f2 * gz * sn;
      }
    }
    for (int pol = 0; pol < 3; ++pol) {
      double x = acc[pol].x, y = acc[pol].y;
      for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off);
        y += __shfl_xor(y, off);
      }
      if (lane == 0) red[wave][pol] = make_double2(x, y);
    }
    __syncthreads();
    if (t < 3) {
      double2 s = make_double2(0., 0.);
      for (int w = 0; w < 4; ++w) {
        s.x += red[w][t].x;
        s.y += red[w][t].y;
      }
      A[3 * a + t] = s;
    }
    __syncthreads();
  }
  const double2* ev = e3 + size_t(iq) * nm * nm;
  for (int nu = t; nu < nm; nu += 256) {
    double2 s = make_double2(0., 0.);
    for (int k = 0; k < 3 * d.natoms; ++k) {
      const double2 a = A[k], e = ev[k + size_t(nm) * nu];
      s.x += a.x * e.x - a.y * e.y;
      s.y += a.x * e.y + a.y * e.x;
    }
    out[size_t(iq) * nm + nu] = make_double2(-s.y, s.x);  // times i
  }
}

void launch_polar_part1(const PolarDev& d, int nq, int nm, const double* q3, const double2* e3, double2* out,
                        hipStream_t s) {
  if (nq <= 0) return;
  PHX_REQUIRE(d.natoms >= 1 && 3 * d.natoms == nm && nm <= kPolarMaxModes, "polar term: nm must be 3 natoms <= 36");
  hipLaunchKernelGGL(k_polar_part1, dim3(nq), dim3(256), 0, s, d, nm, q3, e3, out);
  PHX_HIP(hipGetLastError());
}

void PolarHost::set(const phx_polar* p, hipStream_t s) {
  if (!p) {
    on = false;
    return;
  }
  PHX_REQUIRE(p->natoms >= 1 && p->natoms <= kPolarMaxModes / 3 && p->atomPos && p->bornCharges && p->dielectric &&
                  p->volume > 0.,
              "phx_polar: natoms in 1..12, positions, Born charges, dielectric tensor and volume required");
  for (int i = 0; i < 3; ++i)
    PHX_REQUIRE(p->coarseMesh[i] >= 0 && p->coarseMesh[i] <= 64, "phx_polar: coarse mesh out of range");
  std::vector<double> pos(p->atomPos, p->atomPos + 3 * p->natoms), born(p->bornCharges, p->bornCharges + 9 * p->natoms);
  dPos.alloc(pos.size());
  dBorn.alloc(born.size());
  PHX_HIP(hipMemcpyAsync(dPos.p, pos.data(), dPos.bytes(), hipMemcpyHostToDevice, s));
  PHX_HIP(hipMemcpyAsync(dBorn.p, born.data(), dBorn.bytes(), hipMemcpyHostToDevice, s));
  PHX_HIP(hipStreamSynchronize(s));
  d.natoms = p->natoms;
  d.volume = p->volume;
  for (int i = 0; i < 9; ++i) {
    d.recip[i] = p->recip[i];
    d.eps[i] = p->dielectric[i];
  }
  for (int i = 0; i < 3; ++i) d.coarse[i] = p->coarseMesh[i];
  d.pos = dPos.p;
  d.born = dBorn.p;
  on = true;
}

}  // namespace phx
