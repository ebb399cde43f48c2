// Host side of the phx_elph C ABI: the InteractionElPhWan mirror, the batched
// Bloch diagonalizations and the linewidth builder, all driving the gfx950
// kernels of eig.hip, bloch.hip and elph.hip on one HIP stream per handle.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/phx_elph.h"
#include "coupling_kernel.hpp"
#include "elph_file.hpp"
#include "harmonic_io.hpp"
#include "phel.hpp"
#include "polar.hpp"
#include "phx_common.hpp"
#include "screen.hpp"

namespace phx {

// ---- launchers (eig.hip, bloch.hip, elph.hip)
void launch_jacobi(int N, int M, double2* A, double* W, int* info, const double* masses,
                   hipStream_t stream);
void launch_bloch(bool phonon, int nk, int nw, int nR, const double* kc, const double* R,
                  const double* rw, const void* coef, double2* out, hipStream_t stream);
void launch_ph_long_range(int nq, const double* qc, const LongRange& lr, double2* D,
                          hipStream_t stream);
void launch_vel_operator(bool phonon, int np, int n, double delta, int gammaRule, const double* kc,
                         const double2* Uc, const double2* M7, const double* W7, const double* masses,
                         double2* vel, hipStream_t s);
void launch_vel_degenerate(int np, int n, double threshold, int gammaRule, const double* E,
                           const double* kc, double2* vel, int* tooBig, hipStream_t s);
void launch_synth_gw(double2* gW, size_t n, size_t blk, const double* decay, uint64_t seed,
                     hipStream_t s);
void launch_synth_gw_f32(float2* gW, size_t n, size_t blk, const double* decay, uint64_t seed,
                         hipStream_t s);
void launch_to_float(const double* a, float* b, size_t n, hipStream_t s);
void synth_gw_host(double* gW, size_t n, size_t blk, const double* decay, uint64_t seed);
template <typename T> void launch_fill(T* p, size_t n, T v, hipStream_t s);
void launch_phase_table(int rows, int nR, const double* kc, const double* R, const double* scale,
                        double sign, double2* P, hipStream_t s);
template <typename T>
void launch_phase_rows(int nrows, int rowsPad, int nRp, int Kp, const double* kc, const double* R,
                       T* A, hipStream_t s);
void launch_gather_eig(int na, int nw, const int* ik, const int* lo, const int* nb, const int* off,
                       const double2* U, double2* pool, hipStream_t s);
void launch_transpose_e3(int nk, int nm, const double2* ev, void* out, bool fp32, hipStream_t s);
void launch_scatter_ph(int nq, int nm, const int* iq, const double* w, const double2* ev,
                       double* wAll, double2* evAll, hipStream_t s);


// coupling kernels: CoupParams / SepParams and their launchers (instantiated in
// coupling_f64.hip and coupling_f32.hip)
extern template void launch_coupling<double>(const CoupParams&, bool, hipStream_t);
extern template void launch_coupling<float>(const CoupParams&, bool, hipStream_t);
extern template void launch_separable<double>(const SepParams&, const CoupParams&, bool, int, int,
                                              hipStream_t);
extern template void launch_separable<float>(const SepParams&, const CoupParams&, bool, int, int,
                                             hipStream_t);
extern template void launch_separable_zy<double>(const SepParams&, int, int, hipStream_t, bool);
extern template void launch_separable_zy<float>(const SepParams&, int, int, hipStream_t, bool);
extern template bool launch_separable_x<double>(const SepParams&, const CoupParams&, bool, hipStream_t, hipEvent_t,
                                                  const EpFork&);
extern template bool launch_separable_x<float>(const SepParams&, const CoupParams&, bool, hipStream_t, hipEvent_t,
                                                  const EpFork&);
int coupling_rows_per_tile_f64(int ldb);



// Run f(i) for i in [0, n) on the host's hardware threads.
template <typename F>
static void parallel_for(int n, F&& f) {
  int nt = int(std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 32u));
  nt = std::max(1, std::min(nt, n));
  std::atomic<int> next(0);
  auto work = [&] {
    for (;;) {
      int i0 = next.fetch_add(64);
      if (i0 >= n) break;
      for (int i = i0; i < std::min(n, i0 + 64); ++i) f(i);
    }
  };
  std::vector<std::thread> th;
  for (int i = 1; i < nt; ++i) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
}

// ---- error state
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error() { return g_err.c_str(); }

template <typename F>
static int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -3;
  } catch (...) {
    set_error("unknown error");
    return -3;
  }
}

template <typename T>
static void h2d(DevBuf<T>& d, const T* h, size_t n, hipStream_t s = 0) {
  d.alloc(n);
  if (n) PHX_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
static void h2d(DevBuf<T>& d, const std::vector<T>& v, hipStream_t s = 0) {
  h2d(d, v.data(), v.size(), s);
}
// Same into a grow-only buffer: reallocates only when n exceeds its capacity.
template <typename T>
static void h2d_grow(DevBuf<T>& d, const T* h, size_t n, hipStream_t s) {
  d.reserve(n);
  if (n) PHX_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
static void h2d_grow(DevBuf<T>& d, const std::vector<T>& v, hipStream_t s) {
  h2d_grow(d, v.data(), v.size(), s);
}

// Device copy of the long-range term's data (LongRange holds its pointers).
struct DevLongRange {
  DevBuf<double> tau, born, eps, G, cst;
  LongRange lr;
  bool on() const { return lr.nat > 0 && lr.nG > 0; }
  void upload(int nat, int nG, const double* tau_, const double* born_, const double* eps_,
              const double* G_, const double* cst_, double norm, double gMax, hipStream_t s) {
    lr = LongRange();
    if (nat <= 0 || nG <= 0) return;
    PHX_REQUIRE(tau_ && born_ && eps_ && G_ && cst_, "long-range term: null array");
    PHX_REQUIRE(nat <= 12, "long-range term: at most 12 atoms");
    h2d(tau, tau_, size_t(nat) * 3, s);
    h2d(born, born_, size_t(nat) * 9, s);
    h2d(eps, eps_, 9, s);
    h2d(G, G_, size_t(nG) * 3, s);
    h2d(cst, cst_, size_t(nat) * 9, s);
    lr.nat = nat;
    lr.nG = nG;
    lr.tau = tau.p;
    lr.born = born.p;
    lr.eps = eps.p;
    lr.G = G.p;
    lr.cst = cst.p;
    lr.norm = norm;
    lr.gMax = gMax;
  }
};

// Writes elements [first, first + n) of the flattened coupling tensor, as
// interleaved (re, im) doubles, into a host buffer.
using GwReader = std::function<void(uint64_t first, uint64_t n, double* dst)>;

// Coupling tensor from a host-side reader into device memory: two pinned
// staging buffers, the read of one chunk overlapping the copy of the other;
// FP32 handles (precision != FP64) convert each chunk on the device, so no
// full FP64 copy is ever resident.
static void upload_gw_stream(const GwReader& rd, size_t n, int precision, DevBuf<double2>& gW,
                             DevBuf<float2>& gWf, hipStream_t s) {
  const size_t chunk = std::max<size_t>(1, std::min<size_t>(n, size_t(4) << 20));  // 64 MiB
  struct Staging {
    double2* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Staging() {
      for (int b = 0; b < 2; ++b) {
        if (ev[b]) (void)hipEventSynchronize(ev[b]);
        if (ev[b]) (void)hipEventDestroy(ev[b]);
        if (pin[b]) (void)hipHostFree(pin[b]);
      }
    }
  } st;
  for (int b = 0; b < 2; ++b) {
    PHX_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.pin[b]), chunk * sizeof(double2),
                          hipHostMallocDefault));
    PHX_HIP(hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming));
  }
  DevBuf<double2> stage;
  if (precision == PHX_FP64) {
    gWf.release();
    gW.alloc(n);
  } else {
    gW.release();
    gWf.alloc(n);
    stage.alloc(2 * chunk);
  }
  bool used[2] = {false, false};
  size_t it = 0;
  for (size_t first = 0; first < n; first += chunk, ++it) {
    const int b = int(it & 1);
    const size_t k = std::min(chunk, n - first);
    if (used[b]) PHX_HIP(hipEventSynchronize(st.ev[b]));  // pin[b] and stage[b] are free again
    rd(first, k, reinterpret_cast<double*>(st.pin[b]));
    if (precision == PHX_FP64) {
      PHX_HIP(hipMemcpyAsync(gW.p + first, st.pin[b], k * sizeof(double2), hipMemcpyHostToDevice, s));
    } else {
      double2* sb = stage.p + size_t(b) * chunk;
      PHX_HIP(hipMemcpyAsync(sb, st.pin[b], k * sizeof(double2), hipMemcpyHostToDevice, s));
      launch_to_float(reinterpret_cast<const double*>(sb), reinterpret_cast<float*>(gWf.p + first),
                      2 * k, s);
    }
    PHX_HIP(hipEventRecord(st.ev[b], s));
    used[b] = true;
  }
  PHX_HIP(hipStreamSynchronize(s));
}

// constants.h:9-57 (restated)
static const double kPi = 3.14159265358979323846;
static double phonon_cutoff() {
  const double speedLightSi = 299792458;
  const double hPlanckSi = 6.62607015e-34;
  const double hartreeSi = 4.359744650e-18;
  const double twoPi = 2. * kPi;
  const double teraHertzAu = hPlanckSi / twoPi / hartreeSi * 1.0e+12;
  const double ryToTHz = 1. / teraHertzAu / 2. / twoPi;
  const double ryToCmm1 = 1.0e10 * ryToTHz / speedLightSi;
  return 5. / ryToCmm1;  // el_scattering.cpp:127
}

// Particle::getPopulation (particle.cpp:64-96)
static double population(bool bose, double energy, double temperature, double chemPot) {
  double population = 0.;
  if (temperature <= 0.) {
    if (energy <= chemPot) population = 1.;
    return population;
  }
  double y = (energy - chemPot) / temperature;
  if (bose) {
    population = 1. / (std::exp(y) - 1);
    if (population < 0.) population = 0.;
  } else {
    population = 1. / (std::exp(y) + 1);
    if (population < 0.) population = 0.;
    if (population > 1.) population = 1.;
  }
  return population;
}

// Grouping of consecutive degenerate levels used by symmetrizeCoupling
// (scattering.cpp:1215-1238): start / size of the group of every level.
static void degenerate_groups(const double* e, int n, int* gstart, int* gsize) {
  for (int i = 0; i < n;) {
    int deg = 0;
    for (int j = i; j < n; ++j) {
      if (std::abs(e[i] - e[j]) < 1.0e-6) deg++;
      else break;
    }
    for (int j = 0; j < deg; ++j) {
      gstart[i + j] = i;
      gsize[i + j] = deg;
    }
    i += deg;
  }
}

// Cartesian coordinates of mesh point `index` (points.cpp:222-248).
static void mesh_cart(const double* recip, const int* mesh, int index, double* c) {
  int x, y, z;
  mesh_coords(index, mesh[0], mesh[1], x, y, z);
  double xc[3] = {double(x) / double(mesh[0]), double(y) / double(mesh[1]),
                  double(z) / double(mesh[2])};
  for (int i = 0; i < 3; ++i) c[i] = recip[i] * xc[0] + recip[i + 3] * xc[1] + recip[i + 6] * xc[2];
}

// Integer coordinates n_d = b_d.R / 2pi of lattice vectors R (3 x n, Cartesian),
// folded into [0, m_d).  False when some R is not a lattice vector.
static bool folded_lattice_coords(const double* recip, const int* mesh, const double* R, int n,
                                  int* f) {
  for (int r = 0; r < n; ++r)
    for (int d = 0; d < 3; ++d) {
      double x = 0.;
      for (int i = 0; i < 3; ++i) x += recip[i + 3 * d] * R[3 * r + i];
      x /= 2. * kPi;
      const double v = std::round(x);
      if (!(std::abs(v - x) < 1e-6)) return false;
      f[3 * r + d] = ((int(v) % mesh[d]) + mesh[d]) % mesh[d];
    }
  return true;
}

// e^{2 pi i t / m}
static double2 twiddle(long long t, int m) {
  const double a = 2. * kPi * double(t % m) / double(m);
  return make_double2(std::cos(a), std::sin(a));
}

// ============================================================ batched eig / Bloch
static void zheev_batched(int M, int N, double* A, double* W, int* info) {
  hipStream_t s = 0;
  DevBuf<double2> dA(size_t(M) * N * N);
  DevBuf<double> dW(size_t(M) * N);
  DevBuf<int> dI(M);
  PHX_HIP(hipMemcpy(dA.p, A, dA.bytes(), hipMemcpyHostToDevice));
  launch_jacobi(N, M, dA.p, dW.p, dI.p, nullptr, s);
  PHX_HIP(hipMemcpy(A, dA.p, dA.bytes(), hipMemcpyDeviceToHost));
  PHX_HIP(hipMemcpy(W, dW.p, dW.bytes(), hipMemcpyDeviceToHost));
  std::vector<int> inf(M);
  PHX_HIP(hipMemcpy(inf.data(), dI.p, dI.bytes(), hipMemcpyDeviceToHost));
  if (info) std::copy(inf.begin(), inf.end(), info);
  for (int i = 0; i < M; ++i)
    if (inf[i] < 0) throw Error(-4, "batched zheev: matrix " + std::to_string(i) + " did not converge");
}

static void bloch_diag(bool phonon, int nk, const double* k, int n, int nR, const double* R,
                       const double* rw, const void* coef, const double* masses, int massScaling,
                       double* energies, double* eigvecs, const PhononModel* lrModel = nullptr,
                       int reps = 0, double* msOut = nullptr) {
  hipStream_t s = 0;
  DevBuf<double> dk, dR, drw, dm;
  h2d(dk, k, size_t(nk) * 3);
  h2d(dR, R, size_t(nR) * 3);
  h2d(drw, rw, size_t(nR));
  DevBuf<double2> dcoefc;
  DevBuf<double> dcoefr;
  if (phonon) h2d(dcoefr, static_cast<const double*>(coef), size_t(nR) * n * n);
  else h2d(dcoefc, static_cast<const double2*>(coef), size_t(nR) * n * n);
  DevBuf<double2> H(size_t(nk) * n * n);
  DevBuf<double> W(size_t(nk) * n);
  DevBuf<int> info(nk);
  launch_bloch(phonon, nk, n, nR, dk.p, dR.p, drw.p,
               phonon ? static_cast<const void*>(dcoefr.p) : static_cast<const void*>(dcoefc.p),
               H.p, s);
  DevLongRange dlr;
  if (phonon && lrModel && lrModel->hasDielectric) {
    const PhononModel& m = *lrModel;
    dlr.upload(m.nat, m.nG, m.positions.data(), m.born.data(), m.dielectric, m.gVectors.data(),
               m.lrConstant.data(), m.lrNorm, m.gMax, s);
    if (dlr.on()) launch_ph_long_range(nk, dk.p, dlr.lr, H.p, s);
  }
  const double* mptr = nullptr;
  std::vector<double> ones;
  if (phonon) {
    if (massScaling) {
      h2d(dm, masses, size_t(n));
    } else {
      // hermitize and divide by sqrt(M_m M_n) but keep eigenvectors unscaled:
      // run with the masses and undo the eigenvector scaling afterwards
      h2d(dm, masses, size_t(n));
    }
    mptr = dm.p;
  }
  launch_jacobi(n, nk, H.p, W.p, info.p, mptr, s);
  if (reps > 0 && msOut) {
    // timed repetitions on device-resident inputs: the Bloch build and the
    // eigensolve, each bracketed by its own events
    hipEvent_t ev[3];
    for (auto& e : ev) PHX_HIP(hipEventCreate(&e));
    double tb = 0., tj = 0.;
    for (int r = 0; r < reps; ++r) {
      PHX_HIP(hipEventRecord(ev[0], s));
      launch_bloch(phonon, nk, n, nR, dk.p, dR.p, drw.p,
                   phonon ? static_cast<const void*>(dcoefr.p) : static_cast<const void*>(dcoefc.p), H.p, s);
      PHX_HIP(hipEventRecord(ev[1], s));
      launch_jacobi(n, nk, H.p, W.p, info.p, mptr, s);
      PHX_HIP(hipEventRecord(ev[2], s));
      PHX_HIP(hipEventSynchronize(ev[2]));
      float a = 0.f, b = 0.f;
      PHX_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
      PHX_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
      tb += a;
      tj += b;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    msOut[0] = tb / reps;
    msOut[1] = tj / reps;
  }
  PHX_HIP(hipMemcpy(energies, W.p, W.bytes(), hipMemcpyDeviceToHost));
  PHX_HIP(hipMemcpy(eigvecs, H.p, H.bytes(), hipMemcpyDeviceToHost));
  std::vector<int> inf(nk);
  PHX_HIP(hipMemcpy(inf.data(), info.p, info.bytes(), hipMemcpyDeviceToHost));
  for (int i = 0; i < nk; ++i)
    if (inf[i] < 0) throw Error(-4, "Bloch eigensolve did not converge");
  if (phonon && !massScaling) {
    for (int q = 0; q < nk; ++q)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double sc = std::sqrt(masses[i]);
          double* x = eigvecs + 2 * (size_t(q) * n * n + i + size_t(n) * j);
          x[0] *= sc;
          x[1] *= sc;
        }
  }
}

// Energies, eigenvectors and the velocity operator at np points
// (kokkosBatchedDiagonalizeWithVelocities of ElectronH0Wannier,
// electron_h0_wannier.cpp:888-999, and of PhononH0, phonon_h0_kokkos.cpp:378-532):
// the Hamiltonians at k and k +- 1e-8 e_i in one Bloch launch, the eigensolves,
// the velocity operator and the degenerate-block treatment on the device.
// vel: np x n x n x 3 complex, (p, m, n, i) row-major.  Phonon eigenvectors
// are returned divided by sqrt(M).
static void diag_velocities(bool phonon, int np, const double* k, int n, int nR, const double* R,
                            const double* rw, const void* coef, const double* masses,
                            const PhononModel* lrModel, int gammaRule, double* energies,
                            double* eigvecs, double* vel) {
  PHX_REQUIRE(gammaRule == 0 || gammaRule == 1, "velocities: gamma_rule must be 0 or 1");
  hipStream_t s = 0;
  const double delta = 1.0e-8;
  const double threshold =
      phonon ? 2.0e-5 * phonon_cutoff()          // 1e-4 cm^-1 (phonon_h0_kokkos.cpp:527)
             : 0.000001 / units::energyRyToEv;   // 1 micro-eV (electron_h0_wannier.cpp:897)
  const size_t nn = size_t(n) * n;
  std::vector<double> k7(size_t(np) * 21);
  for (int p = 0; p < np; ++p)
    for (int j = 0; j < 7; ++j) {
      double* d = &k7[(size_t(p) * 7 + j) * 3];
      for (int i = 0; i < 3; ++i) d[i] = k[3 * size_t(p) + i];
      if (j > 0) {
        const int dir = (j - 1) / 2;
        if ((j - 1) % 2 == 0) d[dir] += delta;
        else d[dir] -= delta;
      }
    }
  DevBuf<double> dk7, dkc, dR, drw, dm;
  h2d(dk7, k7, s);
  h2d(dkc, k, size_t(np) * 3, s);
  h2d(dR, R, size_t(nR) * 3, s);
  h2d(drw, rw, size_t(nR), s);
  DevBuf<double2> dcoefc;
  DevBuf<double> dcoefr;
  if (phonon) {
    h2d(dcoefr, static_cast<const double*>(coef), size_t(nR) * nn, s);
    h2d(dm, masses, size_t(n), s);
  } else {
    h2d(dcoefc, static_cast<const double2*>(coef), size_t(nR) * nn, s);
  }
  const int np7 = 7 * np;
  DevBuf<double2> H7(size_t(np7) * nn);
  launch_bloch(phonon, np7, n, nR, dk7.p, dR.p, drw.p,
               phonon ? static_cast<const void*>(dcoefr.p) : static_cast<const void*>(dcoefc.p), H7.p, s);
  DevBuf<double2> Uc(size_t(np) * nn), dvel(size_t(np) * nn * 3);
  DevBuf<double> Wc(size_t(np) * n);
  DevBuf<int> flag(1);
  launch_fill<int>(flag.p, 1, 0, s);
  std::vector<int> inf;
  if (phonon) {
    DevLongRange dlr;
    if (lrModel && lrModel->hasDielectric) {
      const PhononModel& m = *lrModel;
      dlr.upload(m.nat, m.nG, m.positions.data(), m.born.data(), m.dielectric, m.gVectors.data(),
                 m.lrConstant.data(), m.lrNorm, m.gMax, s);
      if (dlr.on()) launch_ph_long_range(np7, dk7.p, dlr.lr, H7.p, s);
    }
    DevBuf<double> W7(size_t(np7) * n);
    DevBuf<int> info(np7);
    launch_jacobi(n, np7, H7.p, W7.p, info.p, dm.p, s);
    launch_vel_operator(true, np, n, delta, gammaRule, dkc.p, nullptr, H7.p, W7.p, dm.p, dvel.p, s);
    PHX_HIP(hipMemcpy2DAsync(Wc.p, size_t(n) * sizeof(double), W7.p, 7 * size_t(n) * sizeof(double),
                             size_t(n) * sizeof(double), np, hipMemcpyDeviceToDevice, s));
    PHX_HIP(hipMemcpy2DAsync(Uc.p, nn * sizeof(double2), H7.p, 7 * nn * sizeof(double2), nn * sizeof(double2),
                             np, hipMemcpyDeviceToDevice, s));
    inf.resize(np7);
    PHX_HIP(hipMemcpyAsync(inf.data(), info.p, info.bytes(), hipMemcpyDeviceToHost, s));
  } else {
    PHX_HIP(hipMemcpy2DAsync(Uc.p, nn * sizeof(double2), H7.p, 7 * nn * sizeof(double2), nn * sizeof(double2),
                             np, hipMemcpyDeviceToDevice, s));
    DevBuf<int> info(np);
    launch_jacobi(n, np, Uc.p, Wc.p, info.p, nullptr, s);
    launch_vel_operator(false, np, n, delta, gammaRule, dkc.p, Uc.p, H7.p, nullptr, nullptr, dvel.p, s);
    inf.resize(np);
    PHX_HIP(hipMemcpyAsync(inf.data(), info.p, info.bytes(), hipMemcpyDeviceToHost, s));
  }
  launch_vel_degenerate(np, n, threshold, gammaRule, Wc.p, dkc.p, dvel.p, flag.p, s);
  int tooBig = 0;
  PHX_HIP(hipMemcpyAsync(&tooBig, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipMemcpyAsync(energies, Wc.p, Wc.bytes(), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipMemcpyAsync(eigvecs, Uc.p, Uc.bytes(), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipMemcpyAsync(vel, dvel.p, dvel.bytes(), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < inf.size(); ++i)
    if (inf[i] < 0) throw Error(-4, "eigensolve did not converge");
  if (tooBig)
    throw Error(-1, "velocities: a degenerate group of " + std::to_string(tooBig) +
                        " bands (at most 8 are supported)");
}

// ============================================================ InteractionElPhWan
struct ElphHandle {
  int nw = 0, nm = 0, nRp = 0, nRe = 0, precision = PHX_FP64;
  hipStream_t stream = nullptr;
  DevBuf<double2> gW;   // FP64 coupling tensor (FP64 handles)
  DevBuf<float2> gWf;   // FP32 copy read by the reduced-precision GEMMs
  DevBuf<double> Re, degE, Rp, degP;
  const void* gw_ptr() const {
    return precision == PHX_FP64 ? static_cast<const void*>(gW.p) : static_cast<const void*>(gWf.p);
  }
  // cached state of the last cacheElPh
  int nb1 = -1;
  int Kp = 0, ldb = 0;
  DevBuf<double> Bd;
  DevBuf<float> Bf;
  DevBuf<double2> U1;
  DevBuf<int> meta;  // U1off(1), nb1(1), panelBase(1), panelKK(nb1), panelB1(nb1), kkNb1(1)
  // per-call API (cacheElPh / calcCouplingSquared): constant tables made on the
  // first call, grow-only scratch reused by later calls
  DevBuf<double> invE, W0;          // 1/deg(R_e); zero phonon energies (store mode)
  DevBuf<double2> phP1;             // 1/deg(R_p) as complex
  DevBuf<double> sK, sQ, sOut, sA;  // k1, q3, |g|^2, FP64 phase rows
  DevBuf<float> sAf;                // FP32 phase rows
  DevBuf<double2> sPhE, sU2, sE3;
  DevBuf<int> sU2off, sNb2, sE3idx;
  DevBuf<long long> sOutOff;
  // polar term: parameters (phx_elph_set_polar) and per-call buffers
  PolarHost polar;
  DevBuf<double2> sPolX;
  ~ElphHandle() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static void elph_common_init(ElphHandle* H, int nw, int nm, int nRp, int nRe, const double* elR,
                             const double* elDeg, const double* phR, const double* phDeg,
                             int precision) {
  PHX_REQUIRE(nw >= 1 && nw <= 32 && nm >= 1 && nm <= 36 && nRp >= 1 && nRe >= 1,
              "phx_elph_create: unsupported sizes");
  PHX_REQUIRE(precision == PHX_FP64 || precision == PHX_FP32,
              "phx_elph_create: precision not available");
  H->nw = nw;
  H->nm = nm;
  H->nRp = nRp;
  H->nRe = nRe;
  H->precision = precision;
  PHX_HIP(hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking));
  h2d(H->Re, elR, size_t(nRe) * 3, H->stream);
  h2d(H->degE, elDeg, size_t(nRe), H->stream);
  h2d(H->Rp, phR, size_t(nRp) * 3, H->stream);
  h2d(H->degP, phDeg, size_t(nRp), H->stream);
  H->Kp = int(round_up(2 * size_t(nRp), 16));
  H->ldb = int(round_up(2 * size_t(nm) * nw, 16));
}

// cacheElPh for one k1 into nb1 panels; phase mode: 0 = 1/dp (explicit q),
// 1 = e^{-ik1.Rp}/dp (mesh split).
static void elph_cache(ElphHandle* H, const double* U1, int nb1, const double* k1) {
  PHX_REQUIRE(nb1 >= 1 && nb1 <= H->nw, "cacheElPh: bad nb1");
  hipStream_t s = H->stream;
  const size_t pstride = size_t(H->Kp) * H->ldb;
  if (nb1 > H->nb1 || H->nb1 < 0) {
    if (H->precision == PHX_FP64) {
      H->Bd.alloc(pstride * nb1);
      launch_fill<double>(H->Bd.p, H->Bd.n, 0., s);
    } else {
      H->Bf.alloc(pstride * nb1);
      launch_fill<float>(H->Bf.p, H->Bf.n, 0.f, s);
    }
  }
  H->nb1 = nb1;
  h2d_grow(H->U1, reinterpret_cast<const double2*>(U1), size_t(H->nw) * nb1, s);
  std::vector<int> meta(4 + 2 * nb1, 0);
  meta[0] = 0;    // U1off
  meta[1] = nb1;  // nb1 / kkNb1
  meta[2] = 0;    // panelBase
  meta[3] = 0;    // panelKK base (unused)
  for (int b = 0; b < nb1; ++b) {
    meta[4 + b] = 0;          // panelKK
    meta[4 + nb1 + b] = b;    // panelB1
  }
  h2d_grow(H->meta, meta, s);
  h2d_grow(H->sK, k1, 3, s);
  // phE = e^{ik1.Re}/de ; phP = 1/dp (explicit-q mode: e^{iq.Rp} sits in the A rows)
  {
    if (!H->invE.p) {
      std::vector<double> hDegE(H->nRe), hDegP(H->nRp);
      PHX_HIP(hipMemcpyAsync(hDegE.data(), H->degE.p, H->degE.bytes(), hipMemcpyDeviceToHost, s));
      PHX_HIP(hipMemcpyAsync(hDegP.data(), H->degP.p, H->degP.bytes(), hipMemcpyDeviceToHost, s));
      PHX_HIP(hipStreamSynchronize(s));
      std::vector<double> invE(H->nRe);
      for (int i = 0; i < H->nRe; ++i) invE[i] = 1. / hDegE[i];
      h2d(H->invE, invE, s);
      std::vector<double2> pp(H->nRp);
      for (int i = 0; i < H->nRp; ++i) pp[i] = make_double2(1. / hDegP[i], 0.);
      h2d(H->phP1, pp, s);
    }
    H->sPhE.reserve(size_t(H->nRe));
    launch_phase_table(1, H->nRe, H->sK.p, H->Re.p, H->invE.p, 1.0, H->sPhE.p, s);
    CacheParams cp{};
    cp.nRe = H->nRe;
    cp.nRp = H->nRp;
    cp.nm = H->nm;
    cp.nw = H->nw;
    cp.nkk = 1;
    cp.gW = H->gw_ptr();
    cp.phE = H->sPhE.p;
    cp.phP = H->phP1.p;
    cp.U1pool = H->U1.p;
    cp.U1off = H->meta.p + 0;
    cp.nb1 = H->meta.p + 1;
    cp.panelBase = H->meta.p + 2;
    cp.panelStride = pstride;
    cp.ldb = H->ldb;
    if (H->precision == PHX_FP64) {
      cp.B = H->Bd.p;
      launch_cache<double, double>(cp, s);
    } else {
      cp.B = H->Bf.p;
      launch_cache<float, float>(cp, s);
    }
    PHX_HIP(hipStreamSynchronize(s));
  }
}

static void elph_coupling_sq(ElphHandle* H, int nq, const double* U2s, const int* nb2s,
                             const double* e3s, const double* q3, const double* polar, double* out) {
  PHX_REQUIRE(H->nb1 >= 1, "calcCouplingSquared called before cacheElPh");
  if (nq <= 0) return;
  hipStream_t s = H->stream;
  const int nw = H->nw, nm = H->nm, nb1 = H->nb1;
  const int rows = coupling_rows_per_tile_f64(H->ldb);
  const int nMtiles = ceil_div(nq, rows);
  const int rowsPad = nMtiles * rows;
  std::vector<int> u2off(nq), nb2v(nq), e3idx(nq);
  std::vector<long long> outoff(nq);
  long long totU2 = 0, totOut = 0;
  int nb2max = 1;
  for (int i = 0; i < nq; ++i) {
    PHX_REQUIRE(nb2s[i] >= 1 && nb2s[i] <= nw, "calcCouplingSquared: bad nb2");
    u2off[i] = int(totU2);
    nb2v[i] = nb2s[i];
    e3idx[i] = i;
    outoff[i] = totOut;
    totU2 += (long long)nw * nb2s[i];
    totOut += (long long)nb1 * nb2s[i] * nm;
    nb2max = std::max(nb2max, nb2s[i]);
  }
  h2d_grow(H->sQ, q3, size_t(nq) * 3, s);
  h2d_grow(H->sU2, reinterpret_cast<const double2*>(U2s), size_t(totU2), s);
  h2d_grow(H->sE3, reinterpret_cast<const double2*>(e3s), size_t(nq) * nm * nm, s);
  h2d_grow(H->sU2off, u2off, s);
  h2d_grow(H->sNb2, nb2v, s);
  h2d_grow(H->sE3idx, e3idx, s);
  h2d_grow(H->sOutOff, outoff, s);
  H->sOut.reserve(size_t(totOut));
  if (!H->W0.p) {
    H->W0.alloc(size_t(nm));
    launch_fill<double>(H->W0.p, H->W0.n, 0., s);
  }
  CoupParams p{};
  p.lda = H->Kp;
  p.Kp = H->Kp;
  p.ldb = H->ldb;
  p.panelStride = size_t(H->Kp) * H->ldb;
  p.M = nq;
  p.nm = nm;
  p.nw = nw;
  p.nPanels = nb1;
  p.nMtiles = nMtiles;
  p.panelKK = H->meta.p + 4;
  p.panelB1 = H->meta.p + 4 + nb1;
  p.U2pool = H->sU2.p;
  p.rowU2off = H->sU2off.p;
  p.rowNb2 = H->sNb2.p;
  p.E3pool = H->sE3.p;
  p.Wpool = H->W0.p;
  p.qmode = 0;
  p.rowE3 = H->sE3idx.p;
  p.out = H->sOut.p;
  p.rowOutOff = H->sOutOff.p;
  p.kkNb1 = H->meta.p + 1;
  p.nb2max = nb2max;
  if (polar) {
    // the reference adds the term only where |q3| > 1e-8 (interaction_elph.cpp:289)
    std::vector<double2> x(size_t(nq) * nm);
    for (int i = 0; i < nq; ++i) {
      const double qn = std::sqrt(q3[3 * i] * q3[3 * i] + q3[3 * i + 1] * q3[3 * i + 1] + q3[3 * i + 2] * q3[3 * i + 2]);
      for (int nu = 0; nu < nm; ++nu)
        x[size_t(i) * nm + nu] = qn > 1.0e-8 ? make_double2(polar[2 * (size_t(i) * nm + nu)],
                                                             polar[2 * (size_t(i) * nm + nu) + 1])
                                             : make_double2(0., 0.);
    }
    h2d_grow(H->sPolX, x, s);
    p.polarX = H->sPolX.p;
    p.U1pool = H->U1.p;
    p.U1off = H->meta.p + 0;
  }
  if (H->precision == PHX_FP64) {
    H->sA.reserve(size_t(rowsPad) * H->Kp);
    launch_phase_rows<double>(nq, rowsPad, H->nRp, H->Kp, H->sQ.p, H->Rp.p, H->sA.p, s);
    p.A = H->sA.p;
    p.B = H->Bd.p;
    launch_coupling<double>(p, false, s);
  } else {
    H->sAf.reserve(size_t(rowsPad) * H->Kp);
    launch_phase_rows<float>(nq, rowsPad, H->nRp, H->Kp, H->sQ.p, H->Rp.p, H->sAf.p, s);
    p.A = H->sAf.p;
    p.B = H->Bf.p;
    launch_coupling<float>(p, false, s);
  }
  PHX_HIP(hipMemcpyAsync(out, H->sOut.p, size_t(totOut) * sizeof(double), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipStreamSynchronize(s));
}

// ============================================================ linewidth builder
// One batch of outer points, prepared on the host: its device metadata lives
// here (one object per batch of a captured sweep), its host copies stay
// alive until the uploads have run.
struct BatchMeta {
  int nkk = 0, nPanels = 0, nMtiles = 0, M = 0, k2b = 0, k2e = 0, nLinesB = 0, nb1Max = 0, nb2max = 1;
  bool line = false;
  double flops = 0.;  // algorithmic FLOPs of the coupling stage (timing)
  // [U1off nkk][nb1 nkk][panelBase nkk][kkMesh nkk][kkStateOff nkk][panelKK P][panelB1 P][panelGStart P][panelGSize P]
  std::vector<int> hmeta, hLineKK;
  std::vector<double> hk1;
  std::vector<double2> hwLine, hphLine;
  std::vector<float2> hphLineF;
  DevBuf<int> meta, lineKK;
  DevBuf<double> k1;
  DevBuf<double2> wLine, phLineD;
  DevBuf<float2> phLineF;
};

struct LwHandle {
  std::vector<double> ownT, ownMu;  // copies of the problem's temperatures and chemical potentials
  // adaptive smearing: group velocities of the states and the scaled reciprocal cell
  int adaptive = 0;
  DevBuf<double> stateV;
  double qT[9] = {};
  phx_problem prob{};
  int precision = PHX_FP64, k1Batch = 16;
  hipStream_t stream = nullptr;
  int nk = 0, na = 0, numStates = 0, nw = 0, nm = 0, nRp = 0, nRe = 0, nCalc = 1;
  // host copies
  std::vector<double> energiesAll;  // nk x nw
  std::vector<int> active, lo, hi, nb, stateOff, u2off, gStart, gSize;
  std::vector<double> stateE;
  std::vector<double> kcart;       // nk x 3
  std::vector<double> omegaAll;    // nk x nm
  std::vector<unsigned char> qInSet;
  // device
  DevBuf<double2> gW;   // FP64 coupling tensor (FP64 handles)
  DevBuf<float2> gWf;   // FP32 copy (reduced-precision handles)
  DevBuf<double> Re, degE, Rp, degP, invDegE, invDegP;
  const void* gw_ptr() const {
    return precision == PHX_FP64 ? static_cast<const void*>(gW.p) : static_cast<const void*>(gWf.p);
  }
  DevBuf<double2> U2pool;
  DevBuf<double> e2pool, fermi, bose, Wall, lw;
  DevBuf<double2> E3all;
  DevBuf<float2> E3Tf;         // separable transform: eigenvectors as (q, nu, nu2), FP32 ...
  DevBuf<double2> E3Td;        // ... or FP64
  DevBuf<int> rowMesh, rowNb2, rowU2off, rowStateOff, gStartD, gSizeD;
  DevBuf<double> kcartActive;  // na x 3
  int Kp = 0, ldb = 0, rows = 32, rowsPad = 0;
  DevBuf<double> Ad;
  DevBuf<float> Af;
  // per-batch
  int maxPanels = 0;
  DevBuf<double> Bd;
  DevBuf<float> Bf;
  DevBuf<double> partial;
  DevBuf<double2> phE, phP;
  BatchMeta cur;                    // the batch being run by phx_lw_run
  std::vector<BatchMeta> plan;      // a prepared sweep (phx_lw_plan)
  hipGraphExec_t graphExec = nullptr;
  long long planPairs = 0;
  // separable mesh transform of the R_ph sum (transform == PHX_TRANSFORM_SEPARABLE)
  int transform = 0;
  int nLines = 0, nG0 = 0, maxLineE = 0, maxGrpE = 0;
  DevBuf<int> lineStart, lineF2, lineRp, grpStart, grpF1, grpLine, g0val, meshToRow;
  DevBuf<double2> twx, twy, twz;
  DevBuf<double> T1d, T2d;
  DevBuf<float> T1f, T2f;
  // cacheElPh on mesh lines: outer points of a run that share (j1, j2) are
  // batched together when the line holds at least lineMin of them (0 = off)
  int lineMin = 0;
  int panelChunk = 0;      // panels per separable-transform launch (bounds T1 / T2 / g3); 0 = automatic
  int splitLast = 1;       // last stage as k_sep_xg + epilogue where they apply (2: the first epilogue,
                           // k_sep_ep; 3: the electron rotation in k_sep_xg, g' rows through HBM)
  DevBuf<float2> G3f;      // their g3 rows of one panel chunk, (slot, jx, col), or g' rows (slot, jx, b2, nu)
  // the split stage's epilogue of chunk i on streamEp, overlapping chunk i + 1's
  // transform stages on `stream`: chunks alternate between two sets of g3 rows
  // and row masks; epEv holds two events per chunk (x-stage done, epilogue done)
  int epOverlap = 0;
  // FP32 split stage: the y-stage inside the x-stage kernel (k_sep_yx), no T2 in HBM
  int fuseYX = 1;
  double yxT1Bytes = 0.;   // timed launches of k_sep_yx: T1 bytes they read
  long long yxLaunches = 0;
  hipStream_t streamEp = nullptr;
  std::vector<hipEvent_t> epEv;
  DevBuf<float2> G3f2;
  DevBuf<unsigned long long> keepMask2;
  int linesPerBatch = kLinesPerBatch;  // mesh lines per line batch (1 .. kLinesPerBatch)
  bool reLattice = false;  // every R_e is a lattice vector (required by the line form)
  int nGe = 0;
  std::vector<int> reF1, reF2, grpF0e;  // n1, n2 mod m of the grouped R_e; n0 mod m0 per group
  std::vector<double> reInvDeg;         // 1/de of the grouped R_e
  DevBuf<int> reOrderD, reGroupEndD;
  int nRePad = 0;                       // groups padded to multiples of 4 entries
  DevBuf<int2> rePadD;                  // (gW row, group-order position), (-1, -1) = padding
  DevBuf<int> reBlkEndD;                // per 4-entry block: group index if it ends its group, else -1
  DevBuf<double2> TgD;
  DevBuf<float2> TgF;

  // learned q-block screen (separable transform with the energy screen on)
  int lsMode = 0;                  // 0 off, 1 apply the trained network
  bool lsCollect = false;          // collecting training samples (no linewidth accumulation)
  double lsThr = 1e-8, lsAudit = 0.01, lsFailFactor = 10., lsSample = 0.02;
  uint64_t lsSeed = 0x5EEDull, lsBatch = 0;
  bool mlpTrained = false;
  DevBuf<float> mlpP, mlpG, mlpM1, mlpM2, mlpScratch, mlpLoss, lsFeat, lsPred, trXd, trTd, trXb, trTb;
  DevBuf<int> lsNscr, lsSlots, lsFail;
  DevBuf<unsigned char> lsMaskC, lsMaskA, lsMaskS;
  DevBuf<double> lsThrAbs, lsPanelStats, lsStats, partialAudit;
  DevBuf<long long> lsCounts;      // [blocks in the energy screen, learned skips, audits, fallbacks]
  std::vector<float> trX, trT;     // collected training samples
  std::vector<int> hNscr, hSel;
  std::vector<unsigned char> hMask;
  double lsLoss[2] = {0., 0.};
  double lsMargin = 0.;            // 99.9th percentile of (target - prediction) on the training set
  // energy-conservation screen of the mesh transform (units of sigma, 0 = off)
  double screen = 0.;
  DevBuf<double> qWlo, qWhi;
  DevBuf<int> qModeGrp;        // per q and mode: first mode of its degenerate group | size << 8
  DevBuf<double2> rowE2range;  // per active point: lowest and highest windowed band energy
  DevBuf<int4> meshRow;        // per mesh point: (active row or -1, nb, U pool offset, state offset)
  DevBuf<unsigned long long> keepMask;  // last-stage row masks of one panel chunk
  DevBuf<double> screenOf;     // per state: screen width (sigma) of the panel whose b1 group starts there
  // polar term of the direct transform: parameters and polarCorrectionPart1 per mesh q (nk x nm)
  PolarHost polar;
  DevBuf<double2> polX;
  DevBuf<double> xminPart;     // per (panel of a chunk, block): closest approach to resonance
  // per-kernel timing with HIP events on `stream` (cache, coupling, collect)
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool evPending = false;
  double kernelMs[3] = {0., 0., 0.};
  long long kernelLaunches = 0;
  double couplingFlops = 0.;  // algorithmic FLOPs of the coupling launches timed
  // last-stage launches of the separable transform: events summed at the next sync,
  // (start, x-stage start, mid, epilogue start, end) per launch; the split stage's
  // x-stage kernel runs between the second and mid, the last two are recorded on
  // the stream its epilogue ran on
  std::vector<hipEvent_t> xev;
  std::vector<char> xevSplit;        // per launch: whether it ran split
  size_t xevUsed = 0;
  double xMs = 0., xgMs = 0., epMs = 0.;
  long long xLaunches = 0, splitLaunches = 0;
  DevBuf<unsigned long long> xKept;  // rows computed by the last stage (timed launches)
  DevBuf<unsigned long long> xBands; // sum of their band counts (split epilogue)
  DevBuf<unsigned long long> xSlots; // slots the split x-stage computed (timed launches)
  ~LwHandle() {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : xev) (void)hipEventDestroy(e);
    for (auto& e : epEv) (void)hipEventDestroy(e);
    if (streamEp) (void)hipStreamDestroy(streamEp);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Whether the separable transform's last stage runs split (x-stage and
// epilogue as two kernels, g3 rows through HBM): FP32 data, nm, nw <= 16,
// nm nw <= 192, row masks (x mesh <= 64); the launcher also needs <= 24
// x-groups (and, with the learned screen, the lean epilogue), and falls back
// to the fused kernel otherwise.
static bool lw_split_last(const LwHandle* L, const phx_problem& p) {
  return L->splitLast && L->transform == PHX_TRANSFORM_SEPARABLE && L->precision == PHX_FP32 && p.nm <= 16 &&
         p.nw <= 16 && p.nm * p.nw <= 192 && p.mesh[0] <= 64;
}

// Panels per separable-transform launch: the handle's setting, or (0) as many
// as keep one chunk's g3 rows of the split last stage within kChunkG3Bytes,
// between 16 and 128: fewer, larger launches have shorter tails (cfg2: 128
// panels, 12.6 GB of g3; a 60^3 mesh with 18 modes: 51).
constexpr double kChunkG3Bytes = 24. * (1ull << 30);
static int lw_panel_chunk(const LwHandle* L, const phx_problem& p, int nPanels, int nb2max) {
  if (L->panelChunk > 0) return std::min(nPanels, L->panelChunk);
  const double ncol = double(p.nm) * p.nw;
  const double perPanel =
      double(p.mesh[1]) * p.mesh[2] * p.mesh[0] * std::max(ncol, 2. * nb2max * p.nm) * sizeof(float2);
  const int c = int(std::min(128., std::max(16., std::floor(kChunkG3Bytes / perPanel))));
  return std::max(1, std::min(nPanels, c));
}

// Fold the events of the last timed batch into kernelMs (stream must be idle).
static void lw_harvest_timing(LwHandle* L) {
  if (!L->evPending) return;
  float a = 0.f, b = 0.f, c = 0.f;
  PHX_HIP(hipEventElapsedTime(&a, L->ev[0], L->ev[1]));
  PHX_HIP(hipEventElapsedTime(&b, L->ev[1], L->ev[2]));
  PHX_HIP(hipEventElapsedTime(&c, L->ev[2], L->ev[3]));
  L->kernelMs[0] += a;
  L->kernelMs[1] += b;
  L->kernelMs[2] += c;
  L->kernelLaunches += 1;
  L->evPending = false;
  for (size_t i = 0; i + 5 <= L->xevUsed; i += 5) {
    float x = 0.f;
    PHX_HIP(hipEventElapsedTime(&x, L->xev[i], L->xev[i + 4]));
    L->xMs += x;
    L->xLaunches += 1;
    if (L->xevSplit[i / 5]) {
      float a = 0.f, b = 0.f;
      PHX_HIP(hipEventElapsedTime(&a, L->xev[i + 1], L->xev[i + 2]));
      PHX_HIP(hipEventElapsedTime(&b, L->xev[i + 3], L->xev[i + 4]));
      L->xgMs += a;
      L->epMs += b;
      L->splitLaunches += 1;
    }
  }
  L->xevUsed = 0;
  L->xevSplit.clear();
}

static void lw_setup(LwHandle* L, const phx_problem* P, int precision, int k1Batch,
                     int transform, const ElphFile* file = nullptr) {
  const phx_problem& p = *P;
  PHX_REQUIRE(precision == PHX_FP64 || precision == PHX_FP32 ||
                  (precision == PHX_BF16X3 && transform == PHX_TRANSFORM_SEPARABLE),
              "phx_lw_setup: precision not available (split-BF16x3 runs with the separable transform)");
  PHX_REQUIRE(p.nw >= 1 && p.nw <= 32 && p.nm >= 1 && p.nm <= 36, "phx_lw_setup: unsupported sizes");
  PHX_REQUIRE(p.nCalc >= 1 && p.nCalc <= kCMax, "phx_lw_setup: 1..4 (T, mu) calculations supported");
  PHX_REQUIRE(p.nw <= 44 && p.nm <= 44, "phx_lw_setup: eigensolver size");
  L->prob = p;
  // the (T, mu) lists are read again after set-up: keep copies
  L->ownT.assign(p.temperatures, p.temperatures + p.nCalc);
  L->ownMu.assign(p.chemPots, p.chemPots + p.nCalc);
  L->prob.temperatures = L->ownT.data();
  L->prob.chemPots = L->ownMu.data();
  L->precision = precision;
  L->k1Batch = std::min(kCacheNKK, std::max(1, k1Batch));
  PHX_REQUIRE(transform == PHX_TRANSFORM_DIRECT || transform == PHX_TRANSFORM_SEPARABLE,
              "phx_lw_setup: unknown transform");
  L->transform = transform;
  L->nw = p.nw;
  L->nm = p.nm;
  L->nRp = p.nRp;
  L->nRe = p.nRe;
  L->nCalc = p.nCalc;
  L->nk = p.mesh[0] * p.mesh[1] * p.mesh[2];
  PHX_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
  hipStream_t s = L->stream;
  const int nk = L->nk, nw = p.nw, nm = p.nm;

  // ---- electrons on the full mesh (ActiveBandStructure::buildAsPostprocessing,
  //      active_bandstructure.cpp:765-1075)
  L->kcart.resize(3 * size_t(nk));
  for (int ik = 0; ik < nk; ++ik) mesh_cart(p.recip, p.mesh, ik, &L->kcart[3 * size_t(ik)]);
  DevBuf<double> dk, dRh, dDegh;
  DevBuf<double2> dh0R;
  h2d(dk, L->kcart, s);
  h2d(dRh, p.Rh, size_t(p.nRh) * 3, s);
  h2d(dDegh, p.degh, size_t(p.nRh), s);
  h2d(dh0R, reinterpret_cast<const double2*>(p.h0R), size_t(p.nRh) * nw * nw, s);
  DevBuf<double2> Uall(size_t(nk) * nw * nw);
  DevBuf<double> Eall(size_t(nk) * nw);
  DevBuf<int> info(nk);
  launch_bloch(false, nk, nw, p.nRh, dk.p, dRh.p, dDegh.p, dh0R.p, Uall.p, s);
  launch_jacobi(nw, nk, Uall.p, Eall.p, info.p, nullptr, s);
  L->energiesAll.resize(size_t(nk) * nw);
  std::vector<int> inf(nk);
  PHX_HIP(hipMemcpyAsync(L->energiesAll.data(), Eall.p, Eall.bytes(), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipMemcpyAsync(inf.data(), info.p, info.bytes(), hipMemcpyDeviceToHost, s));
  PHX_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < nk; ++i)
    PHX_REQUIRE(inf[i] >= 0, "electron eigensolve did not converge");

  // ---- energy window (Window::internalEnWindow, window.cpp:137-157), ascending mesh order
  for (int ik = 0; ik < nk; ++ik) {
    int l = -1, h = -1;
    for (int ib = 0; ib < nw; ++ib) {
      double e = L->energiesAll[size_t(ik) * nw + ib];
      if (e < p.winMax && e > p.winMin) {
        if (l < 0) l = ib;
        h = ib;
      }
    }
    if (l < 0) continue;
    L->active.push_back(ik);
    L->lo.push_back(l);
    L->hi.push_back(h);
  }
  L->na = int(L->active.size());
  const int na = L->na;
  PHX_REQUIRE(na > 0, "energy window selects no states");
  int st = 0, uo = 0;
  for (int a = 0; a < na; ++a) {
    int n = L->hi[a] - L->lo[a] + 1;
    L->nb.push_back(n);
    L->stateOff.push_back(st);
    L->u2off.push_back(uo);
    for (int b = 0; b < n; ++b) L->stateE.push_back(L->energiesAll[size_t(L->active[a]) * nw + L->lo[a] + b]);
    st += n;
    uo += n * nw;
  }
  L->numStates = st;
  L->gStart.resize(st);
  L->gSize.resize(st);
  for (int a = 0; a < na; ++a)
    degenerate_groups(&L->stateE[L->stateOff[a]], L->nb[a], &L->gStart[L->stateOff[a]],
                      &L->gSize[L->stateOff[a]]);
  for (int i = 0; i < st; ++i) PHX_REQUIRE(L->gSize[i] <= kGMax, "degenerate group too large");
  {
    DevBuf<int> dAct, dLo, dNb, dOff;
    h2d(dAct, L->active, s);
    h2d(dLo, L->lo, s);
    h2d(dNb, L->nb, s);
    h2d(dOff, L->u2off, s);
    L->U2pool.alloc(size_t(uo));
    launch_gather_eig(na, nw, dAct.p, dLo.p, dNb.p, dOff.p, Uall.p, L->U2pool.p, s);
    PHX_HIP(hipStreamSynchronize(s));
  }
  h2d(L->rowMesh, L->active, s);
  h2d(L->rowNb2, L->nb, s);
  h2d(L->rowU2off, L->u2off, s);
  h2d(L->rowStateOff, L->stateOff, s);
  h2d(L->e2pool, L->stateE, s);
  // ---- group velocities for adaptive smearing: the single-point velocity
  //      operator of every active point (buildAsPostprocessing with velocities,
  //      active_bandstructure.cpp:1032-1071), the window's bands kept
  PHX_REQUIRE(p.smearing == PHX_SMEAR_GAUSSIAN || p.smearing == PHX_SMEAR_ADAPTIVE,
              "phx_problem: unknown smearing");
  if (p.smearing == PHX_SMEAR_ADAPTIVE) {
    PHX_REQUIRE(nw <= 44, "adaptive smearing: at most 44 bands");
    std::vector<double> sv(3 * size_t(st), 0.);
    const int chunk = 2048;
    std::vector<double> kc, E, U, V;
    for (int a0 = 0; a0 < na; a0 += chunk) {
      const int n = std::min(chunk, na - a0);
      kc.resize(3 * size_t(n));
      for (int a = 0; a < n; ++a)
        for (int i = 0; i < 3; ++i) kc[3 * size_t(a) + i] = L->kcart[3 * size_t(L->active[a0 + a]) + i];
      E.resize(size_t(n) * nw);
      U.resize(2 * size_t(n) * nw * nw);
      V.resize(2 * size_t(n) * nw * nw * 3);
      diag_velocities(false, n, kc.data(), nw, p.nRh, p.Rh, p.degh, p.h0R, nullptr, nullptr,
                      PHX_GAMMA_POINT, E.data(), U.data(), V.data());
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < L->nb[a0 + a]; ++b) {
          const int bf = L->lo[a0 + a] + b;
          for (int i = 0; i < 3; ++i)
            sv[3 * size_t(L->stateOff[a0 + a] + b) + i] =
                V[2 * ((((size_t(a) * nw + bf) * nw) + bf) * 3 + i)];
        }
    }
    h2d(L->stateV, sv, s);
    // qTensor: the reciprocal cell with row i divided by mesh(i) (delta_function.cpp:58-62)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) L->qT[3 * i + j] = p.recip[i + 3 * j] / p.mesh[i];
    L->adaptive = 1;
  }
  {
    std::vector<double2> er(na);
    for (int a = 0; a < na; ++a)
      er[a] = make_double2(L->stateE[L->stateOff[a]], L->stateE[L->stateOff[a] + L->nb[a] - 1]);
    h2d(L->rowE2range, er, s);
    PHX_HIP(hipStreamSynchronize(s));
  }
  h2d(L->gStartD, L->gStart, s);
  h2d(L->gSizeD, L->gSize, s);
  {
    std::vector<double> kca(3 * size_t(na));
    for (int a = 0; a < na; ++a)
      for (int i = 0; i < 3; ++i) kca[3 * a + i] = L->kcart[3 * size_t(L->active[a]) + i];
    h2d(L->kcartActive, kca, s);
  }
  Uall.release();

  // ---- q-set: every folded k2 - k1 of active points (helper_el_scattering.cpp:95-124)
  L->qInSet.assign(nk, 0);
  {
    std::vector<char> occ(nk, 0);
    for (int a = 0; a < na; ++a) occ[L->active[a]] = 1;
    const int m0 = p.mesh[0], m1 = p.mesh[1], m2 = p.mesh[2];
    parallel_for(nk, [&](int q) {
      int qx, qy, qz;
      mesh_coords(q, m0, m1, qx, qy, qz);
      for (int a = 0; a < na; ++a) {
        int x, y, z;
        mesh_coords(L->active[a], m0, m1, x, y, z);
        int i2 = ((z + qz) % m2) * m0 * m1 + ((y + qy) % m1) * m0 + (x + qx) % m0;
        if (occ[i2]) {
          L->qInSet[q] = 1;
          break;
        }
      }
    });
  }
  std::vector<int> qlist;
  for (int q = 0; q < nk; ++q)
    if (L->qInSet[q]) qlist.push_back(q);
  const int nq = int(qlist.size());

  // ---- phonons on the q-set (PhononH0 Kokkos path, phonon_h0_kokkos.cpp:6-219)
  L->Wall.alloc(size_t(nk) * nm);
  L->E3all.alloc(size_t(nk) * nm * nm);
  launch_fill<double>(L->Wall.p, L->Wall.n, 0., s);
  launch_fill<double>(reinterpret_cast<double*>(L->E3all.p), 2 * L->E3all.n, 0., s);
  {
    std::vector<double> qc(3 * size_t(nq));
    for (int j = 0; j < nq; ++j)
      for (int i = 0; i < 3; ++i) qc[3 * j + i] = L->kcart[3 * size_t(qlist[j]) + i];
    DevBuf<double> dq, dRf, dwf, dmat, dmass;
    DevBuf<int> dql;
    h2d(dq, qc, s);
    h2d(dql, qlist, s);
    h2d(dRf, p.Rf, size_t(p.nRf) * 3, s);
    h2d(dwf, p.wf, size_t(p.nRf), s);
    h2d(dmat, p.mat2R, size_t(p.nRf) * nm * nm, s);
    h2d(dmass, p.masses, size_t(nm), s);
    DevBuf<double2> D(size_t(nq) * nm * nm);
    DevBuf<double> w(size_t(nq) * nm);
    DevBuf<int> qinfo(nq);
    launch_bloch(true, nq, nm, p.nRf, dq.p, dRf.p, dwf.p, dmat.p, D.p, s);
    DevLongRange dlr;
    if (p.nG > 0) {
      PHX_REQUIRE(3 * p.natoms == nm, "phx_problem: natoms * 3 != nm with a long-range term");
      dlr.upload(p.natoms, p.nG, p.atomPos, p.bornCharges, p.dielectric, p.gVectors, p.lrConstant,
                 p.lrNorm, p.gMax, s);
      launch_ph_long_range(nq, dq.p, dlr.lr, D.p, s);
    }
    launch_jacobi(nm, nq, D.p, w.p, qinfo.p, dmass.p, s);
    launch_scatter_ph(nq, nm, dql.p, w.p, D.p, L->Wall.p, L->E3all.p, s);
    if (L->transform == PHX_TRANSFORM_SEPARABLE) {
      const bool f32 = L->precision != PHX_FP64;
      if (f32) L->E3Tf.alloc(size_t(nk) * nm * nm);
      else L->E3Td.alloc(size_t(nk) * nm * nm);
      launch_transpose_e3(nk, nm, L->E3all.p, f32 ? static_cast<void*>(L->E3Tf.p) : static_cast<void*>(L->E3Td.p),
                          f32, s);
    }
    std::vector<int> qinf(nq);
    PHX_HIP(hipMemcpyAsync(qinf.data(), qinfo.p, qinfo.bytes(), hipMemcpyDeviceToHost, s));
    L->omegaAll.resize(size_t(nk) * nm);
    PHX_HIP(hipMemcpyAsync(L->omegaAll.data(), L->Wall.p, L->Wall.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nq; ++j) PHX_REQUIRE(qinf[j] >= 0, "phonon eigensolve did not converge");
  }
  // lowest / highest phonon energy above the cutoff at every q (screen tables)
  {
    const double cut = phonon_cutoff();
    std::vector<double> lo(nk, 1e300), hi(nk, -1e300);
    for (int q : qlist)
      for (int nu = 0; nu < nm; ++nu) {
        const double w = L->omegaAll[size_t(q) * nm + nu];
        if (w >= cut) {
          lo[q] = std::min(lo[q], w);
          hi[q] = std::max(hi[q], w);
        }
      }
    h2d(L->qWlo, lo, s);
    h2d(L->qWhi, hi, s);
    // degenerate mode groups of every q, as symmetrizeCoupling forms them
    // (scattering.cpp:1286-1316): consecutive modes within 1e-6 of the first
    std::vector<int> grp(size_t(nk) * nm, 0);
    for (int q = 0; q < nk; ++q) {
      const double* w = &L->omegaAll[size_t(q) * nm];
      for (int s0 = 0; s0 < nm;) {
        int deg = 0;
        for (int i = s0; i < nm; ++i) {
          if (std::abs(w[s0] - w[i]) < 1.0e-6) ++deg;
          else break;
        }
        for (int i = s0; i < s0 + deg; ++i) grp[size_t(q) * nm + i] = s0 | (deg << 8);
        s0 += deg;
      }
    }
    h2d(L->qModeGrp, grp, s);
    PHX_HIP(hipStreamSynchronize(s));
  }
  // Bose (helper_el_scattering.cpp:261-267) and inner Fermi (el_scattering.cpp:87-109)
  {
    std::vector<double> bose(size_t(nk) * nm * p.nCalc, 0.);
    for (int q : qlist)
      for (int nu = 0; nu < nm; ++nu)
        for (int c = 0; c < p.nCalc; ++c)
          bose[(size_t(q) * nm + nu) * p.nCalc + c] =
              population(true, L->omegaAll[size_t(q) * nm + nu], p.temperatures[c], 0.);
    h2d(L->bose, bose, s);
    std::vector<double> fermi(size_t(p.nCalc) * st);
    for (int c = 0; c < p.nCalc; ++c)
      for (int i = 0; i < st; ++i)
        fermi[size_t(c) * st + i] = population(false, L->stateE[i], p.temperatures[c], p.chemPots[c]);
    h2d(L->fermi, fermi, s);
  }

  // ---- coupling tensor and R vectors
  h2d(L->Re, p.Re, size_t(p.nRe) * 3, s);
  h2d(L->Rp, p.Rp, size_t(p.nRp) * 3, s);
  {
    std::vector<double> ie(p.nRe), ip(p.nRp);
    for (int i = 0; i < p.nRe; ++i) ie[i] = 1. / p.degE[i];
    for (int i = 0; i < p.nRp; ++i) ip[i] = 1. / p.degP[i];
    h2d(L->invDegE, ie, s);
    h2d(L->invDegP, ip, s);
  }
  const size_t gwN = size_t(p.nRe) * p.nRp * nm * nw * nw;
  if (file) {
    upload_gw_stream([file](uint64_t f0, uint64_t k, double* d) { file->read_gw(f0, k, d); }, gwN,
                     precision, L->gW, L->gWf, s);
  } else if (p.gW) {
    h2d(L->gW, reinterpret_cast<const double2*>(p.gW), gwN, s);
    if (precision != PHX_FP64) {
      L->gWf.alloc(gwN);
      launch_to_float(reinterpret_cast<const double*>(L->gW.p), reinterpret_cast<float*>(L->gWf.p),
                      2 * gwN, s);
      PHX_HIP(hipStreamSynchronize(s));
      L->gW.release();
    }
  } else {
    PHX_REQUIRE(p.gwDecay != nullptr, "phx_problem: gW or gwDecay required");
    DevBuf<double> dec;
    h2d(dec, p.gwDecay, size_t(p.nRe) * p.nRp, s);
    if (precision == PHX_FP64) {
      L->gW.alloc(gwN);
      launch_synth_gw(L->gW.p, gwN, size_t(nm) * nw * nw, dec.p, p.gwSeed, s);
    } else {
      L->gWf.alloc(gwN);
      launch_synth_gw_f32(L->gWf.p, gwN, size_t(nm) * nw * nw, dec.p, p.gwSeed, s);
    }
    PHX_HIP(hipStreamSynchronize(s));
  }

  // ---- GEMM A operand: e^{ik2.Rp} rows for every active point
  L->Kp = int(round_up(2 * size_t(p.nRp), 16));
  L->ldb = int(round_up(2 * size_t(nm) * nw, 16));
  L->rows = coupling_rows_per_tile_f64(L->ldb);
  L->rowsPad = int(round_up(size_t(na), L->rows)) + L->rows;
  if (L->transform == PHX_TRANSFORM_DIRECT) {
    if (precision == PHX_FP64) {
      L->Ad.alloc(size_t(L->rowsPad) * L->Kp);
      launch_phase_rows<double>(na, L->rowsPad, p.nRp, L->Kp, L->kcartActive.p, L->Rp.p, L->Ad.p, s);
    } else {
      L->Af.alloc(size_t(L->rowsPad) * L->Kp);
      launch_phase_rows<float>(na, L->rowsPad, p.nRp, L->Kp, L->kcartActive.p, L->Rp.p, L->Af.p, s);
    }
  } else {
    // ---- separable transform: integer coordinates n_d = b_d.R / 2pi of every R_ph,
    //      folded mod m_d; z-lines = distinct (f0, f1), x-groups = distinct f0.
    const int m[3] = {p.mesh[0], p.mesh[1], p.mesh[2]};
    std::vector<int> f(3 * size_t(p.nRp));
    PHX_REQUIRE(folded_lattice_coords(p.recip, p.mesh, p.Rp, p.nRp, f.data()),
                "separable transform: R_ph is not a lattice vector");
    std::vector<int> lineOf(size_t(m[0]) * m[1], -1), grpOf(m[0], -1);
    std::vector<std::vector<int>> lineE;  // R_ph indices of each line
    std::vector<int> lineKey;             // f0 + m0 * f1
    for (int r = 0; r < p.nRp; ++r) {
      const int key = f[3 * r] + m[0] * f[3 * r + 1];
      if (lineOf[key] < 0) {
        lineOf[key] = int(lineE.size());
        lineE.emplace_back();
        lineKey.push_back(key);
      }
      lineE[lineOf[key]].push_back(r);
    }
    std::vector<std::vector<int>> grpE;   // lines of each group
    std::vector<int> g0v;
    for (size_t l = 0; l < lineE.size(); ++l) {
      const int f0 = lineKey[l] % m[0];
      if (grpOf[f0] < 0) {
        grpOf[f0] = int(grpE.size());
        grpE.emplace_back();
        g0v.push_back(f0);
      }
      grpE[grpOf[f0]].push_back(int(l));
    }
    std::vector<int> ls{0}, lf2, lrp, gs{0}, gf1, gl;
    for (auto& e : lineE) {
      for (int r : e) {
        lf2.push_back(f[3 * r + 2]);
        lrp.push_back(r);
      }
      ls.push_back(int(lrp.size()));
      L->maxLineE = std::max(L->maxLineE, int(e.size()));
    }
    for (auto& e : grpE) {
      for (int l : e) {
        gf1.push_back(lineKey[l] / m[0]);
        gl.push_back(l);
      }
      gs.push_back(int(gl.size()));
      L->maxGrpE = std::max(L->maxGrpE, int(e.size()));
    }
    L->nLines = int(lineE.size());
    L->nG0 = int(grpE.size());
    {
      int nbMax = 0;
      for (int n : L->nb) nbMax = std::max(nbMax, n);
      PHX_REQUIRE(nm <= kSepNm && nbMax <= kSepNb2,
                  "phx_lw_setup: the separable transform supports nm <= 32 and at most 8 windowed "
                  "bands per point (use the direct transform)");
      // the z / y stages stage one line / group of panel columns in LDS
      const size_t ncol = size_t(nm) * nw;
      PHX_REQUIRE(size_t(L->maxLineE) * ncol * 16 <= 160 * 1024 && size_t(L->maxGrpE) * ncol * 16 <= 160 * 1024 &&
                      (ncol <= 384 || (size_t(L->nG0) * ncol * (precision == PHX_FP64 ? 16 : 8) <= 96 * 1024)),
                  "phx_lw_setup: panels too wide for the separable transform (use the direct transform)");
    }
    h2d(L->lineStart, ls, s);
    h2d(L->lineF2, lf2, s);
    h2d(L->lineRp, lrp, s);
    h2d(L->grpStart, gs, s);
    h2d(L->grpF1, gf1, s);
    h2d(L->grpLine, gl, s);
    h2d(L->g0val, g0v, s);
    auto tw = [&](DevBuf<double2>& d, int n) {
      std::vector<double2> v(n);
      for (int t = 0; t < n; ++t) v[t] = make_double2(std::cos(2. * kPi * t / n), std::sin(2. * kPi * t / n));
      h2d(d, v, s);
      PHX_HIP(hipStreamSynchronize(s));
    };
    tw(L->twx, m[0]);
    tw(L->twy, m[1]);
    tw(L->twz, m[2]);
    std::vector<int> m2r(nk, -1);
    for (int a = 0; a < na; ++a) m2r[L->active[a]] = a;
    h2d(L->meshToRow, m2r, s);
    std::vector<int4> mr(nk, make_int4(-1, 0, 0, 0));
    for (int a = 0; a < na; ++a) mr[L->active[a]] = make_int4(a, L->nb[a], L->u2off[a], L->stateOff[a]);
    h2d(L->meshRow, mr, s);
    PHX_HIP(hipStreamSynchronize(s));
  }
  // ---- R_e grouped by n0 mod m0 for the mesh-line form of cacheElPh
  {
    std::vector<int> f(3 * size_t(p.nRe));
    L->reLattice = folded_lattice_coords(p.recip, p.mesh, p.Re, p.nRe, f.data());
    if (L->reLattice) {
      std::vector<int> order(p.nRe);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return f[3 * a] < f[3 * b]; });
      std::vector<int> groupEnd(p.nRe, -1);
      L->reF1.resize(p.nRe);
      L->reF2.resize(p.nRe);
      L->reInvDeg.resize(p.nRe);
      L->grpF0e.clear();
      for (int i = 0; i < p.nRe; ++i) {
        const int r = order[i];
        L->reF1[i] = f[3 * r + 1];
        L->reF2[i] = f[3 * r + 2];
        L->reInvDeg[i] = 1. / p.degE[r];
        if (i + 1 == p.nRe || f[3 * order[i + 1]] != f[3 * r]) {
          groupEnd[i] = int(L->grpF0e.size());
          L->grpF0e.push_back(f[3 * r]);
        }
      }
      L->nGe = int(L->grpF0e.size());
      h2d(L->reOrderD, order, s);
      h2d(L->reGroupEndD, groupEnd, s);
      std::vector<int2> pad;
      std::vector<int> blkEnd;
      for (int i = 0, start = 0; i < p.nRe; ++i) {
        pad.push_back(make_int2(order[i], i));
        if (groupEnd[i] >= 0) {
          for (int k = i + 1 - start; k % 4; ++k) pad.push_back(make_int2(-1, -1));
          blkEnd.resize(pad.size() / 4, -1);
          blkEnd.back() = groupEnd[i];
          start = i + 1;
        }
      }
      L->nRePad = int(pad.size());
      h2d(L->rePadD, pad, s);
      h2d(L->reBlkEndD, blkEnd, s);
      PHX_HIP(hipStreamSynchronize(s));
    }
  }
  L->lw.alloc(size_t(p.nCalc) * st);
  launch_fill<double>(L->lw.p, L->lw.n, 0., s);
  PHX_HIP(hipStreamSynchronize(s));
}

// Learned q-block screen for one panel chunk of the separable transform (the
// z / y stages have run).  Collect mode: the exact last stage on a random
// sample of the blocks inside the energy screen, whose features and log
// contributions are appended to the training set (host).  Apply mode:
// features -> MLP -> skip / audit / compute masks; exact stage on the computed
// blocks, exact stage on the audited blocks into a side buffer, audit, and a
// gated exact stage on every skipped block when the audit fails.
template <typename XStage>
static void learned_screen_chunk(LwHandle* L, const CoupParams& cq, XStage&& xstage, int panelBase,
                                 int nPan, int nLinesYZ, int k2b, int k2e) {
  const phx_problem& p = L->prob;
  hipStream_t s = L->stream;
  const int slots = nPan * nLinesYZ;
  L->lsFeat.reserve(size_t(slots) * kMlpIn);
  L->lsNscr.reserve(slots);
  FeatParams fp{};
  fp.nPanels = nPan;
  fp.m0 = p.mesh[0];
  fp.m1 = p.mesh[1];
  fp.m2 = p.mesh[2];
  fp.panelKK = cq.panelKK;
  fp.panelB1 = cq.panelB1;
  fp.panelGStart = cq.panelGStart;
  fp.panelGSize = cq.panelGSize;
  fp.kkMesh = cq.kkMesh;
  fp.kkStateOff = cq.kkStateOff;
  fp.epool = L->e2pool.p;
  fp.meshToRow = L->meshToRow.p;
  fp.rowLo = k2b;
  fp.rowHi = k2e;
  fp.rowE2range = L->rowE2range.p;
  fp.qWlo = L->qWlo.p;
  fp.qWhi = L->qWhi.p;
  fp.mu = p.chemPots[0];
  fp.escale = 0.5 * (p.winMax - p.winMin);
  fp.wscale = 0.;
  for (double w : L->omegaAll) fp.wscale = std::max(fp.wscale, w);
  if (!(fp.wscale > 0.)) fp.wscale = 1.;
  fp.sigma = p.sigma;
  fp.screen = L->screen;
  fp.feat = L->lsFeat.p;
  fp.nscreen = L->lsNscr.p;
  launch_block_features(fp, s);
  L->lsMaskC.reserve(slots);
  if (L->lsCollect) {
    // training sample: blocks inside the energy screen drawn with probability lsSample
    L->hNscr.resize(slots);
    PHX_HIP(hipMemcpyAsync(L->hNscr.data(), L->lsNscr.p, slots * sizeof(int), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    L->hMask.assign(slots, 0);
    L->hSel.clear();
    // a random fraction of every block inside the energy screen, resonant or
    // not: the network predicts the panels' totals from all of them
    for (int i = 0; i < slots; ++i)
      if ((L->hNscr[i] & 0xffff) > 0 &&
          screen_draw(L->lsSeed ^ 0xC011EC7ull, L->lsBatch, panelBase + i / nLinesYZ, nLinesYZ,
                      i % nLinesYZ) < L->lsSample) {
        L->hMask[i] = 1;
        L->hSel.push_back(i);
      }
    const int nSel = int(L->hSel.size());
    PHX_HIP(hipMemcpyAsync(L->lsMaskC.p, L->hMask.data(), slots, hipMemcpyHostToDevice, s));
    CoupParams c1 = cq;
    c1.slotOn = L->lsMaskC.p;
    c1.slotZero = 1;
    xstage(c1);
    if (nSel == 0) return;
    L->lsSlots.reserve(nSel);
    PHX_HIP(hipMemcpyAsync(L->lsSlots.p, L->hSel.data(), nSel * sizeof(int), hipMemcpyHostToDevice, s));
    L->trXb.reserve(size_t(nSel) * kMlpIn);
    L->trTb.reserve(nSel);
    launch_gather_targets(nSel, L->lsSlots.p, cq.partial, L->lsFeat.p, L->trXb.p, L->trTb.p, s);
    const size_t o = L->trT.size();
    L->trX.resize(L->trX.size() + size_t(nSel) * kMlpIn);
    L->trT.resize(o + nSel);
    PHX_HIP(hipMemcpyAsync(L->trX.data() + o * kMlpIn, L->trXb.p, size_t(nSel) * kMlpIn * sizeof(float),
                           hipMemcpyDeviceToHost, s));
    PHX_HIP(hipMemcpyAsync(L->trT.data() + o, L->trTb.p, nSel * sizeof(float), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    return;
  }
  // inference and decisions
  L->lsPred.reserve(slots);
  L->mlpScratch.reserve(mlp_scratch_floats(slots));
  mlp_forward(L->mlpP.p, L->lsFeat.p, slots, L->mlpScratch.p, L->lsPred.p, s);
  L->lsMaskA.reserve(slots);
  L->lsMaskS.reserve(slots);
  L->lsThrAbs.reserve(nPan);
  L->lsPanelStats.reserve(2 * size_t(nPan));
  DecideParams dp{};
  dp.nPanels = nPan;
  dp.nLinesYZ = nLinesYZ;
  dp.panelBase = panelBase;
  dp.pred = L->lsPred.p;
  dp.nscreen = L->lsNscr.p;
  dp.thr = L->lsThr;
  dp.auditFrac = L->lsAudit;
  dp.seed = L->lsSeed;
  dp.batchId = L->lsBatch;
  dp.margin = L->lsMargin;
  dp.feat = L->lsFeat.p;
  dp.computeMask = L->lsMaskC.p;
  dp.auditMask = L->lsMaskA.p;
  dp.skipMask = L->lsMaskS.p;
  dp.thrAbs = L->lsThrAbs.p;
  dp.counts = L->lsCounts.p;
  launch_screen_decide(dp, s);
  CoupParams c1 = cq;
  c1.slotOn = L->lsMaskC.p;
  c1.slotZero = 1;
  xstage(c1);
  L->partialAudit.reserve(size_t(slots) * kGMax * kCMax);
  CoupParams c2 = cq;
  c2.slotOn = L->lsMaskA.p;
  c2.slotZero = 1;
  c2.partial = L->partialAudit.p;
  xstage(c2);
  AuditParams ap{};
  ap.nPanels = nPan;
  ap.nLinesYZ = nLinesYZ;
  ap.auditMask = L->lsMaskA.p;
  ap.thrAbs = L->lsThrAbs.p;
  ap.partialAudit = L->partialAudit.p;
  ap.partial = cq.partial;
  ap.failFactor = L->lsFailFactor;
  ap.fail = L->lsFail.p;
  ap.stats = L->lsStats.p;
  launch_screen_audit(ap, L->lsPanelStats.p, s);
  CoupParams c3 = cq;
  c3.slotOn = L->lsMaskS.p;
  c3.slotZero = 0;
  c3.gate = L->lsFail.p;
  xstage(c3);
  launch_fail_fold(L->lsFail.p, L->lsCounts.p + 3, s);
}

// Host side of one batch of outer points: metadata and line tables into b
// (uploaded on the handle's stream), every scratch buffer the launch needs
// sized.  line: the outer points lie on at most kLinesPerBatch mesh lines and
// cacheElPh runs in its line form.
static void lw_prepare(LwHandle* L, const int* k1s, int nkk, int k2b, int k2e, bool line, BatchMeta& b) {
  const phx_problem& p = L->prob;
  hipStream_t s = L->stream;
  const int nw = L->nw, nm = L->nm;
  // panels of this batch
  int nPanels = 0;
  for (int i = 0; i < nkk; ++i) nPanels += L->nb[k1s[i]];
  const size_t pstride = size_t(L->Kp) * L->ldb;
  if (nPanels > L->maxPanels) {
    L->maxPanels = std::max(nPanels, L->k1Batch * nw);
    if (L->precision == PHX_FP64) {
      L->Bd.alloc(pstride * L->maxPanels);
      launch_fill<double>(L->Bd.p, L->Bd.n, 0., s);
    } else {
      L->Bf.alloc(pstride * L->maxPanels);
      launch_fill<float>(L->Bf.p, L->Bf.n, 0.f, s);
    }
  }
  const int M = k2e - k2b;
  const int nMtiles = L->transform == PHX_TRANSFORM_DIRECT ? ceil_div(M, L->rows)
                                                         : p.mesh[1] * p.mesh[2];
  b.nkk = nkk;
  b.nPanels = nPanels;
  b.nMtiles = nMtiles;
  b.M = M;
  b.k2b = k2b;
  b.k2e = k2e;
  b.line = line;
  L->partial.reserve(size_t(L->maxPanels) * std::max(nMtiles, 1) * kGMax * kCMax);
  // metadata: [U1off nkk][nb1 nkk][panelBase nkk][kkMesh nkk][kkStateOff nkk]
  //           [panelKK P][panelB1 P][panelGStart P][panelGSize P]
  std::vector<int>& m = b.hmeta;
  m.assign(5 * size_t(nkk) + 4 * size_t(nPanels), 0);
  int* U1off = m.data();
  int* nb1 = U1off + nkk;
  int* pbase = nb1 + nkk;
  int* kkMesh = pbase + nkk;
  int* kkState = kkMesh + nkk;
  int* pKK = kkState + nkk;
  int* pB1 = pKK + nPanels;
  int* pGS = pB1 + nPanels;
  int* pGZ = pGS + nPanels;
  std::vector<double>& k1c = b.hk1;
  k1c.assign(3 * size_t(nkk), 0.);
  int pan = 0;
  for (int i = 0; i < nkk; ++i) {
    const int a = k1s[i];
    U1off[i] = L->u2off[a];
    nb1[i] = L->nb[a];
    pbase[i] = pan;
    kkMesh[i] = L->active[a];
    kkState[i] = L->stateOff[a];
    for (int d = 0; d < 3; ++d) k1c[3 * i + d] = L->kcart[3 * size_t(L->active[a]) + d];
    for (int b = 0; b < L->nb[a]; ++b, ++pan) {
      pKK[pan] = i;
      pB1[pan] = b;
      pGS[pan] = L->gStart[L->stateOff[a] + b];
      pGZ[pan] = L->gSize[L->stateOff[a] + b];
    }
  }
  b.meta.reserve(m.size());
  PHX_HIP(hipMemcpyAsync(b.meta.p, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice, s));
  b.k1.reserve(k1c.size());
  PHX_HIP(hipMemcpyAsync(b.k1.p, k1c.data(), k1c.size() * sizeof(double), hipMemcpyHostToDevice, s));
  // line tables (line form) or the phase table size (GEMM form)
  int nLinesB = 0;
  if (line) {
    // lines of the batch: consecutive outer points with equal (j1, j2)
    const int m0 = p.mesh[0], m1 = p.mesh[1], m2 = p.mesh[2];
    std::vector<int>& lk = b.hLineKK;
    lk.clear();
    for (int i = 0; i < nkk; ++i)
      if (i == 0 || L->active[k1s[i]] / m0 != L->active[k1s[i - 1]] / m0) lk.push_back(i);
    nLinesB = int(lk.size());
    lk.push_back(nkk);
    PHX_REQUIRE(nLinesB <= kLinesPerBatch, "line batch: too many mesh lines");
    // rows beyond nLinesB (the line count rounded up for stage a) stay zero
    const int nLinesA = line_count_padded(nLinesB);
    b.hphLine.assign(size_t(nLinesA) * L->nRe, make_double2(0., 0.));  // (R_e, line)
    for (int l = 0; l < nLinesB; ++l) {
      int j0, j1, j2;
      mesh_coords(L->active[k1s[lk[l]]], m0, m1, j0, j1, j2);
      for (int i = 0; i < L->nRe; ++i) {
        const double2 ty = twiddle((long long)j1 * L->reF1[i], m1);
        const double2 tz = twiddle((long long)j2 * L->reF2[i], m2);
        const double d = L->reInvDeg[i];
        b.hphLine[size_t(i) * nLinesA + l] =
            make_double2((ty.x * tz.x - ty.y * tz.y) * d, (ty.x * tz.y + ty.y * tz.x) * d);
      }
    }
    b.hwLine.resize(size_t(nkk) * L->nGe);
    for (int i = 0; i < nkk; ++i) {
      int x, y, z;
      mesh_coords(L->active[k1s[i]], m0, m1, x, y, z);
      for (int g = 0; g < L->nGe; ++g)
        b.hwLine[size_t(i) * L->nGe + g] = twiddle((long long)x * L->grpF0e[g], m0);
    }
    b.lineKK.reserve(lk.size());
    PHX_HIP(hipMemcpyAsync(b.lineKK.p, lk.data(), lk.size() * sizeof(int), hipMemcpyHostToDevice, s));
    b.wLine.reserve(b.hwLine.size());
    PHX_HIP(hipMemcpyAsync(b.wLine.p, b.hwLine.data(), b.hwLine.size() * sizeof(double2),
                           hipMemcpyHostToDevice, s));
    const size_t nph = b.hphLine.size();
    if (L->precision == PHX_FP64) {
      b.phLineD.reserve(nph);
      PHX_HIP(hipMemcpyAsync(b.phLineD.p, b.hphLine.data(), nph * sizeof(double2),
                             hipMemcpyHostToDevice, s));
    } else {
      b.hphLineF.resize(nph);
      for (size_t i = 0; i < nph; ++i)
        b.hphLineF[i] = make_float2(float(b.hphLine[i].x), float(b.hphLine[i].y));
      b.phLineF.reserve(nph);
      PHX_HIP(hipMemcpyAsync(b.phLineF.p, b.hphLineF.data(), nph * sizeof(float2),
                             hipMemcpyHostToDevice, s));
    }
    const size_t N = size_t(L->nRp) * nm * nw * nw;
    const size_t nA = size_t(line_count_padded(nLinesB));
    if (L->precision == PHX_FP64) L->TgD.reserve(nA * L->nGe * N);
    else L->TgF.reserve(nA * L->nGe * N);
  } else {
    L->phE.reserve(size_t(nkk) * L->nRe);
  }
  b.nLinesB = nLinesB;
  L->phP.reserve(size_t(nkk) * L->nRp);
  b.nb2max = 1;
  for (int a = k2b; a < k2e; ++a) b.nb2max = std::max(b.nb2max, L->nb[a]);
  b.nb1Max = 0;
  for (int i = 0; i < nkk; ++i) b.nb1Max = std::max(b.nb1Max, L->nb[k1s[i]]);
  // scratch of the separable transform: one panel chunk of stage buffers, row masks,
  // and the learned screen's per-block arrays
  if (L->transform == PHX_TRANSFORM_SEPARABLE && M > 0) {
    // sized for the handle's full chunk, not this batch's panel count, so that a
    // later, larger batch does not reallocate them
    const int chunk = lw_panel_chunk(L, p, std::numeric_limits<int>::max(), b.nb2max);
    const size_t ncol = size_t(nm) * nw;
    const size_t t1 = size_t(L->nLines) * p.mesh[2] * ncol * 2, t2 = size_t(L->nG0) * p.mesh[1] * p.mesh[2] * ncol * 2;
    if (L->precision == PHX_FP64) {
      L->T1d.reserve(t1 * chunk);
      L->T2d.reserve(t2 * chunk);
    } else {
      L->T1f.reserve(t1 * chunk);
      L->T2f.reserve(t2 * chunk);
    }
    const size_t slots = size_t(chunk) * nMtiles;
    if (p.mesh[0] <= 64) L->keepMask.reserve(slots);
    if (L->screen > 0.) {
      L->screenOf.reserve(size_t(L->numStates));
      L->xminPart.reserve(size_t(chunk) * kXminBlocks);
    }
    if (lw_split_last(L, p)) {  // g3 (complex FP32) or g' (complex FP64, nb2max nm per row)
      const size_t g3 = slots * p.mesh[0] * std::max(ncol, size_t(2) * b.nb2max * nm);
      L->G3f.reserve(g3);
      if (L->epOverlap && p.mesh[0] <= 64) {
        L->G3f2.reserve(g3);
        L->keepMask2.reserve(slots);
      }
    }
    if (L->lsMode && L->mlpTrained) {
      L->lsFeat.reserve(slots * kMlpIn);
      L->lsNscr.reserve(slots);
      L->lsMaskC.reserve(slots);
      L->lsMaskA.reserve(slots);
      L->lsMaskS.reserve(slots);
      L->lsPred.reserve(slots);
      L->mlpScratch.reserve(mlp_scratch_floats(int(slots)));
      L->lsThrAbs.reserve(chunk);
      L->lsPanelStats.reserve(2 * size_t(chunk));
      L->partialAudit.reserve(slots * kGMax * kCMax);
    }
  }
  // algorithmic FLOPs of the coupling stage (per (pair, b1): transform, rotations, |g|^2)
  double f = 0.;
  if (L->transform == PHX_TRANSFORM_DIRECT) {
    for (int a = k2b; a < k2e; ++a) {
      const double nb2 = L->nb[a];
      f += 8. * L->nRp * nm * nw + 8. * nm * nw * nb2 + 8. * nm * nm * nb2 + 3. * nm * nb2;
    }
  } else {
    const double ncol = double(nm) * nw, mx = p.mesh[0], my = p.mesh[1], mz = p.mesh[2];
    f = 8. * ncol * (double(L->nRp) * mz + double(L->nLines) * my * mz + double(L->nG0) * mx * my * mz);
    for (int a = k2b; a < k2e; ++a) {
      const double nb2 = L->nb[a];
      f += 8. * nm * nw * nb2 + 8. * nm * nm * nb2 + 3. * nm * nb2;
    }
  }
  b.flops = f * nPanels;
}

// Device side of one prepared batch: every launch on the handle's stream, no
// allocation, upload or host synchronization (so that a sweep can be captured
// into a graph).
static long long lw_launch(LwHandle* L, const BatchMeta& b) {
  const phx_problem& p = L->prob;
  hipStream_t s = L->stream;
  const int nw = L->nw, nm = L->nm;
  const int nkk = b.nkk, nPanels = b.nPanels, nMtiles = b.nMtiles, M = b.M, k2b = b.k2b, k2e = b.k2e;
  const int nLinesB = b.nLinesB;
  const bool line = b.line;
  const size_t pstride = size_t(L->Kp) * L->ldb;
  const int* dm = b.meta.p;
  const int* dU1off = dm;
  const int* dNb1 = dm + nkk;
  const int* dPbase = dm + 2 * nkk;
  const int* dKKMesh = dm + 3 * nkk;
  const int* dKKState = dm + 4 * nkk;
  const int* dPKK = dm + 5 * nkk;
  const int* dPB1 = dPKK + nPanels;
  const int* dPGS = dPB1 + nPanels;
  const int* dPGZ = dPGS + nPanels;
  // phases: e^{ik1.Re}/de (GEMM form), e^{-ik1.Rp}/dp
  if (!line) launch_phase_table(nkk, L->nRe, b.k1.p, L->Re.p, L->invDegE.p, 1.0, L->phE.p, s);
  launch_phase_table(nkk, L->nRp, b.k1.p, L->Rp.p, L->invDegP.p, -1.0, L->phP.p, s);
  if (L->timing) PHX_HIP(hipEventRecord(L->ev[0], s));
  // cacheElPh for the whole batch
  CacheParams cp{};
  cp.nRe = L->nRe;
  cp.nRp = L->nRp;
  cp.nm = nm;
  cp.nw = nw;
  cp.nkk = nkk;
  cp.gW = L->gw_ptr();
  cp.phE = L->phE.p;
  cp.phP = L->phP.p;
  cp.U1pool = L->U2pool.p;
  cp.U1off = dU1off;
  cp.nb1 = dNb1;
  cp.panelBase = dPbase;
  cp.panelStride = pstride;
  cp.ldb = L->ldb;
  if (line) {
    cp.N = size_t(L->nRp) * nm * nw * nw;
    cp.nGe = L->nGe;
    cp.reOrder = L->reOrderD.p;
    cp.reGroupEnd = L->reGroupEndD.p;
    cp.nRePad = L->nRePad;
    cp.rePad = L->rePadD.p;
    cp.reBlkEnd = L->reBlkEndD.p;
    cp.nLinesB = nLinesB;
    cp.nLinesA = line_count_padded(nLinesB);
    cp.lineKK = b.lineKK.p;
    cp.wLine = b.wLine.p;
  }
  if (L->precision == PHX_FP64) {
    cp.B = L->Bd.p;
    if (line) {
      cp.phLine = b.phLineD.p;
      cp.Tg = L->TgD.p;
      launch_cache_line<double, double>(cp, s);
    } else {
      launch_cache<double, double>(cp, s);
    }
  } else {
    cp.B = L->Bf.p;
    if (line) {
      cp.phLine = b.phLineF.p;
      cp.Tg = L->TgF.p;
      launch_cache_line<float, float>(cp, s);
    } else {
      launch_cache<float, float>(cp, s);
    }
  }
  if (M <= 0) {
    if (L->timing) {
      PHX_HIP(hipEventRecord(L->ev[1], s));
      PHX_HIP(hipEventRecord(L->ev[2], s));
      PHX_HIP(hipEventRecord(L->ev[3], s));
      L->evPending = true;
    }
    return 0;
  }
  if (L->timing) PHX_HIP(hipEventRecord(L->ev[1], s));
  // coupling + fused reduction
  CoupParams cq{};
  cq.lda = L->Kp;
  cq.Kp = L->Kp;
  cq.ldb = L->ldb;
  cq.panelStride = pstride;
  cq.M = M;
  cq.nm = nm;
  cq.nw = nw;
  cq.nPanels = nPanels;
  cq.nMtiles = nMtiles;
  cq.panelKK = dPKK;
  cq.panelB1 = dPB1;
  cq.U2pool = L->U2pool.p;
  cq.rowU2off = L->rowU2off.p + k2b;
  cq.rowNb2 = L->rowNb2.p + k2b;
  cq.E3pool = L->E3all.p;
  cq.E3T = L->precision == PHX_FP64 ? static_cast<const void*>(L->E3Td.p) : static_cast<const void*>(L->E3Tf.p);
  cq.Wpool = L->Wall.p;
  cq.qmode = 1;
  cq.rowMesh = L->rowMesh.p + k2b;
  cq.kkMesh = dKKMesh;
  cq.m0 = p.mesh[0];
  cq.m1 = p.mesh[1];
  cq.m2 = p.mesh[2];
  cq.e2pool = L->e2pool.p;
  cq.rowStateOff = L->rowStateOff.p + k2b;
  cq.e1pool = L->e2pool.p;
  cq.kkStateOff = dKKState;
  cq.panelGStart = dPGS;
  cq.panelGSize = dPGZ;
  cq.fermi = L->fermi.p;
  cq.numStates = L->numStates;
  cq.bose = L->bose.p;
  cq.nCalc = L->nCalc;
  cq.invSigma = L->adaptive ? 0. : 1. / p.sigma;
  cq.deltaPref = L->adaptive ? 0. : 1. / p.sigma / std::sqrt(kPi);
  cq.adaptive = L->adaptive;
  cq.stateV = L->stateV.p;
  for (int i = 0; i < 9; ++i) cq.qT[i] = L->qT[i];
  cq.minWidth = 0.0001 / units::energyRyToEv;  // delta_function.h:107
  cq.norm = 1. / (double(p.mesh[0]) * p.mesh[1] * p.mesh[2]);
  cq.cutoff = phonon_cutoff();
  cq.partial = L->partial.p;
  cq.nb2max = b.nb2max;
  if (L->polar.on) {
    cq.polarX = L->polX.p;
    cq.U1pool = L->U2pool.p;
    cq.U1off = dU1off;
  }
  if (L->transform == PHX_TRANSFORM_DIRECT) {
    if (L->precision == PHX_FP64) {
      cq.A = L->Ad.p + size_t(k2b) * L->Kp;
      cq.B = L->Bd.p;
      launch_coupling<double>(cq, true, s);
    } else {
      cq.A = L->Af.p + size_t(k2b) * L->Kp;
      cq.B = L->Bf.p;
      launch_coupling<float>(cq, true, s);
    }
  } else {
    const int ncol = nm * nw, mx = p.mesh[0], my = p.mesh[1], mz = p.mesh[2];
    SepParams sp{};
    sp.panelStride = pstride;
    sp.ldb = L->ldb;
    sp.ncol = ncol;
    sp.nPanels = nPanels;
    sp.mx = mx;
    sp.my = my;
    sp.mz = mz;
    sp.nLines = L->nLines;
    sp.lineStart = L->lineStart.p;
    sp.lineF2 = L->lineF2.p;
    sp.lineRp = L->lineRp.p;
    sp.twz = L->twz.p;
    sp.t1PanelStride = size_t(L->nLines) * mz * ncol * 2;
    sp.nG0 = L->nG0;
    sp.grpStart = L->grpStart.p;
    sp.grpF1 = L->grpF1.p;
    sp.grpLine = L->grpLine.p;
    sp.twy = L->twy.p;
    sp.t2PanelStride = size_t(L->nG0) * my * mz * ncol * 2;
    // rows are addressed by active index: no k2 offset, range through rowLo/rowHi
    cq.rowU2off = L->rowU2off.p;
    cq.rowNb2 = L->rowNb2.p;
    cq.rowStateOff = L->rowStateOff.p;
    cq.rowLo = k2b;
    cq.rowHi = k2e;
    cq.meshToRow = L->meshToRow.p;
    cq.t2PanelStride = sp.t2PanelStride;
    cq.nG0 = L->nG0;
    cq.g0val = L->g0val.p;
    cq.tw0 = L->twx.p;
    cq.screen = L->screen;
    if (L->screen > 0.) {
      cq.screenOf = L->screenOf.p;
      cq.xminPart = L->xminPart.p;
    }
    cq.qWlo = L->qWlo.p;
    cq.qWhi = L->qWhi.p;
    cq.qModeGrp = L->qModeGrp.p;
    cq.rowE2range = L->rowE2range.p;
    cq.bf16x3 = L->precision == PHX_BF16X3;
    cq.meshRow = L->meshRow.p;
    if (p.mesh[0] <= 64) cq.keepMask = L->keepMask.p;
    if (lw_split_last(L, p)) {
      cq.G3 = L->G3f.p;
      cq.epFirst = L->splitLast == 2;
      cq.epRot = L->splitLast == 3;
    }
    // panels in chunks: the stage buffers T1 / T2 hold one chunk
    const int chunk = lw_panel_chunk(L, p, nPanels, b.nb2max);
    if (L->precision == PHX_FP64) {
      sp.T1 = L->T1d.p;
      sp.T2 = L->T2d.p;
      cq.T2 = L->T2d.p;
    } else {
      sp.T1 = L->T1f.p;
      sp.T2 = L->T2f.p;
      cq.T2 = L->T2f.p;
    }
    double* const partial0 = cq.partial;
    // the epilogue of each chunk on streamEp (the learned screen's passes stay on s)
    const bool learned = L->lsCollect || (L->lsMode && L->mlpTrained);
    // g3 / g' rows of one chunk (what the batch's preparation reserved, unless a
    // setting changed since: then the fused kernel runs, or the chunks run in turn)
    const size_t slotsC = size_t(chunk) * nMtiles;
    const size_t g3Need = slotsC * mx * std::max(size_t(ncol), size_t(2) * b.nb2max * nm);
    const bool splitOk = lw_split_last(L, p) && L->G3f.n >= g3Need;
    const bool overlap = L->epOverlap && !learned && splitOk && L->G3f2.n >= g3Need &&
                         L->keepMask2.n >= slotsC;
    // the y-stage inside the x-stage kernel: FP32 split stage without slot masks
    const bool yx = L->fuseYX && splitOk && !learned && L->precision == PHX_FP32 && L->splitLast != 3 &&
                    L->nG0 <= 24 && sep_yx_lds_bytes(L->nLines, L->nG0, mx) <= 160 * 1024;
    cq.fuseYX = yx;
    const int nChunks = (nPanels + chunk - 1) / chunk;
    if (overlap) {
      if (!L->streamEp) PHX_HIP(hipStreamCreateWithFlags(&L->streamEp, hipStreamNonBlocking));
      while (L->epEv.size() < 2 * size_t(nChunks)) {
        hipEvent_t e;
        PHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        L->epEv.push_back(e);
      }
    }
    bool epPending[2] = {false, false};  // an epilogue on streamEp reads set 0 / 1
    int lastEp = -1;                     // chunk of the last epilogue on streamEp
    for (int p0 = 0, ci = 0; p0 < nPanels; p0 += chunk, ++ci) {
      const int n = std::min(chunk, nPanels - p0);
      const int set = overlap ? (ci & 1) : 0;
      cq.G3 = splitOk ? (set ? L->G3f2.p : L->G3f.p) : nullptr;
      if (p.mesh[0] <= 64) cq.keepMask = set ? L->keepMask2.p : L->keepMask.p;
      sp.nPanels = n;
      cq.nPanels = n;
      cq.panelKK = dPKK + p0;
      cq.panelB1 = dPB1 + p0;
      cq.panelGStart = dPGS + p0;
      cq.panelGSize = dPGZ + p0;
      cq.gsizeMax = *std::max_element(b.hmeta.begin() + 5 * nkk + 3 * nPanels + p0,
                                      b.hmeta.begin() + 5 * nkk + 3 * nPanels + p0 + n);
      cq.partial = partial0 + size_t(p0) * nMtiles * kGMax * kCMax;
      const bool f64 = L->precision == PHX_FP64;
      if (f64) sp.B = L->Bd.p + size_t(p0) * pstride;
      else sp.B = L->Bf.p + size_t(p0) * pstride;
      EpFork fork{};
      if (overlap) {
        fork.ep = L->streamEp;
        fork.xgDone = L->epEv[2 * ci];
      }
      auto xstage = [&](const CoupParams& c0) {
        CoupParams c = c0;
        hipEvent_t mid = nullptr;
        if (L->timing) {
          while (L->xev.size() < L->xevUsed + 5) {
            hipEvent_t e;
            PHX_HIP(hipEventCreate(&e));
            L->xev.push_back(e);
          }
          c.keptRows = L->xKept.p;
          c.keptBands = L->xBands.p;
          c.keptSlots = L->xSlots.p;
          fork.xgStart = L->xev[L->xevUsed + 1];
          mid = L->xev[L->xevUsed + 2];
          fork.epStart = L->xev[L->xevUsed + 3];
          PHX_HIP(hipEventRecord(L->xev[L->xevUsed], s));
        }
        const bool split = f64 ? launch_separable_x<double>(sp, c, true, s, mid, fork)
                               : launch_separable_x<float>(sp, c, true, s, mid, fork);
        const bool forked = split && fork.ep;
        if (L->timing && split && c.fuseYX) {
          L->yxT1Bytes += double(sp.nPanels) * mz * L->nLines * ncol * 8.;
          L->yxLaunches += 1;
        }
        if (forked) {
          PHX_HIP(hipEventRecord(L->epEv[2 * ci + 1], L->streamEp));
          epPending[set] = true;
          lastEp = ci;
        }
        if (L->timing) {
          PHX_HIP(hipEventRecord(L->xev[L->xevUsed + 4], forked ? L->streamEp : s));
          L->xevUsed += 5;
          L->xevSplit.push_back(split ? 1 : 0);
        }
      };
      if (f64) launch_separable_zy<double>(sp, L->maxLineE, L->maxGrpE, s, true);
      else launch_separable_zy<float>(sp, L->maxLineE, L->maxGrpE, s, !yx);
      if (!learned) {
        // this chunk rewrites set `set`: the epilogue of chunk ci - 2 must have read it
        if (epPending[set]) PHX_HIP(hipStreamWaitEvent(s, L->epEv[2 * (ci - 2) + 1], 0));
        epPending[set] = false;
        xstage(cq);
        continue;
      }
      learned_screen_chunk(L, cq, xstage, p0, n, nMtiles, k2b, k2e);
    }
    // the collect (and the next batch) read what the last epilogues wrote
    if (lastEp >= 0) PHX_HIP(hipStreamWaitEvent(s, L->epEv[2 * lastEp + 1], 0));
  }
  if (L->timing) {
    PHX_HIP(hipEventRecord(L->ev[2], s));
    L->couplingFlops += b.flops;
  }
  if (L->lsCollect || (L->lsMode && L->mlpTrained)) ++L->lsBatch;
  if (L->lsCollect) {  // training samples only: the linewidths are left untouched
    if (L->timing) {
      PHX_HIP(hipEventRecord(L->ev[3], s));
      L->evPending = true;
    }
    return (long long)nkk * M;
  }
  CollectParams cl{};
  cl.nkk = nkk;
  cl.nMtiles = nMtiles;
  cl.nCalc = L->nCalc;
  cl.numStates = L->numStates;
  cl.nb1Max = b.nb1Max;
  cl.kkNb1 = dNb1;
  cl.kkPanelBase = dPbase;
  cl.kkStateOff = dKKState;
  cl.panelGStart = dPGS;
  cl.panelGSize = dPGZ;
  cl.partial = L->partial.p;
  cl.linewidth = L->lw.p;
  launch_collect(cl, s);
  if (L->timing) {
    PHX_HIP(hipEventRecord(L->ev[3], s));
    L->evPending = true;
  }
  return (long long)nkk * M;
}


// Prepare and enqueue one batch (phx_lw_run's path: the caller waits for the
// previous batch before its metadata object is reused).
static long long lw_batch(LwHandle* L, const int* k1s, int nkk, int k2b, int k2e, bool line) {
  lw_prepare(L, k1s, nkk, k2b, k2e, line, L->cur);
  return lw_launch(L, L->cur);
}

struct BatchSpec {
  std::vector<int> k1;
  bool line;
};

// The batches of a sweep over the outer points k1list (all when null): whole
// mesh lines with at least lineMin points go into line batches of up to
// linesPerBatch lines (kLinePointsMax points); the rest into GEMM batches of k1Batch.
static std::vector<BatchSpec> lw_batches(LwHandle* L, int nk1, const int* k1list) {
  std::vector<int> all;
  if (!k1list) {
    all.resize(L->na);
    std::iota(all.begin(), all.end(), 0);
    k1list = all.data();
    nk1 = L->na;
  }
  for (int i = 0; i < nk1; ++i) PHX_REQUIRE(k1list[i] >= 0 && k1list[i] < L->na, "phx_lw_run: bad k1");
  std::vector<BatchSpec> out;
  std::vector<int> rest;
  if (L->lineMin > 0) {
    const int m0 = L->prob.mesh[0];
    std::vector<int> order(k1list, k1list + nk1);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return L->active[a] / m0 < L->active[b] / m0;  // mesh index / m0 = line (j1, j2)
    });
    std::vector<int> batch;
    int linesIn = 0;
    auto flush = [&] {
      if (!batch.empty()) out.push_back(BatchSpec{batch, true});
      batch.clear();
      linesIn = 0;
    };
    for (size_t i = 0; i < order.size();) {
      size_t j = i;
      while (j < order.size() && L->active[order[j]] / m0 == L->active[order[i]] / m0) ++j;
      if (int(j - i) >= L->lineMin) {
        for (size_t c = i; c < j; c += kLinePointsMax) {
          const size_t n = std::min<size_t>(kLinePointsMax, j - c);
          if (linesIn == L->linesPerBatch || batch.size() + n > size_t(kLinePointsMax)) flush();
          batch.insert(batch.end(), order.begin() + c, order.begin() + c + n);
          ++linesIn;
        }
      } else {
        rest.insert(rest.end(), order.begin() + i, order.begin() + j);
      }
      i = j;
    }
    flush();
    k1list = rest.data();
    nk1 = int(rest.size());
  }
  for (int i = 0; i < nk1; i += L->k1Batch)
    out.push_back(BatchSpec{std::vector<int>(k1list + i, k1list + std::min(nk1, i + L->k1Batch)), false});
  return out;
}

static long long lw_run(LwHandle* L, int nk1, const int* k1list, int k2b, int k2e) {
  if (k2e < 0) k2e = L->na;
  PHX_REQUIRE(k2b >= 0 && k2b <= k2e && k2e <= L->na, "phx_lw_run: bad k2 range");
  long long pairs = 0;
  for (const BatchSpec& b : lw_batches(L, nk1, k1list)) {
    // the metadata object is reused per batch: wait for the previous batch
    PHX_HIP(hipStreamSynchronize(L->stream));
    lw_harvest_timing(L);
    pairs += lw_batch(L, b.k1.data(), int(b.k1.size()), k2b, k2e, b.line);
  }
  return pairs;
}

// A whole sweep prepared up front (one metadata object per batch).
static void lw_plan(LwHandle* L, int nk1, const int* k1list, int k2b, int k2e) {
  if (k2e < 0) k2e = L->na;
  PHX_REQUIRE(k2b >= 0 && k2b <= k2e && k2e <= L->na, "phx_lw_plan: bad k2 range");
  PHX_REQUIRE(!L->lsCollect, "phx_lw_plan: not while collecting screen samples");
  if (L->graphExec) {
    PHX_HIP(hipGraphExecDestroy(L->graphExec));
    L->graphExec = nullptr;
  }
  PHX_HIP(hipStreamSynchronize(L->stream));
  L->plan.clear();
  L->planPairs = 0;
  const std::vector<BatchSpec> specs = lw_batches(L, nk1, k1list);
  L->plan.resize(specs.size());
  for (size_t i = 0; i < specs.size(); ++i) {
    lw_prepare(L, specs[i].k1.data(), int(specs[i].k1.size()), k2b, k2e, specs[i].line, L->plan[i]);
    L->planPairs += (long long)specs[i].k1.size() * (k2e - k2b);
  }
  PHX_HIP(hipStreamSynchronize(L->stream));  // uploads and buffer fills done
}

// Enqueue the prepared sweep; with graph != 0 it is captured into a hipGraph on
// first use and replayed as one graph launch.  Asynchronous on the handle's stream.
static long long lw_plan_run(LwHandle* L, int graph) {
  PHX_REQUIRE(!L->plan.empty() || L->planPairs == 0, "phx_lw_plan_run: no plan");
  if (!graph) {
    for (const BatchMeta& b : L->plan) lw_launch(L, b);
    return L->planPairs;
  }
  if (!L->graphExec) {
    const bool timing = L->timing;
    L->timing = false;
    hipGraph_t g = nullptr;
    PHX_HIP(hipStreamBeginCapture(L->stream, hipStreamCaptureModeRelaxed));
    try {
      for (const BatchMeta& b : L->plan) lw_launch(L, b);
    } catch (...) {
      (void)hipStreamEndCapture(L->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      L->timing = timing;
      throw;
    }
    PHX_HIP(hipStreamEndCapture(L->stream, &g));
    L->timing = timing;
    hipError_t e = hipGraphInstantiate(&L->graphExec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PHX_HIP(e);
  }
  PHX_HIP(hipGraphLaunch(L->graphExec, L->stream));
  return L->planPairs;
}


// ============================================================ PhElScatteringMatrix
// Phonon linewidths from el-ph scattering (PhElScatteringMatrix::builder,
// phel_scattering.cpp:43-363).  Outer loop over the active electron points k1
// of the k mesh (window as the electron builder's); inner "loop" over every
// phonon wavevector q3 of the q mesh at once: H(k1 + q3) eigensolves (all
// bands), cacheElPh + calcCouplingSquared with the e^{iq3.Rp} rows computed
// once, then the symmetrize + rate kernels of phel.hip.
struct PhelHandle {
  ElphHandle* H = nullptr;  // coupling (not owned)
  hipStream_t s = nullptr;
  int nw = 0, nm = 0, nk = 0, nq = 0, nCalc = 0, na = 0, adaptive = 0, folded = 0, nb1max = 0;
  int qmesh[3] = {0, 0, 0}, kmesh[3] = {0, 0, 0};
  double norm = 0., invSigma = 0., deltaPref = 0., qT[9] = {};
  std::vector<double> kcart, qcart, omega;   // nk x 3, nq x 3, nq x nm
  std::vector<int> active, lo, nb;
  std::vector<double> stateE, U1, v1;          // per active point: energies, (nw x nb) eigvecs, nb x 3
  std::vector<size_t> stateOff, u1Off;
  std::vector<double> T, mu;
  std::vector<double> hRh, hdegh, hh0R;        // electron Hamiltonian on the host (adaptive k2 velocities)
  DevBuf<double> Rh, degh, invT, W3, lw, E2, k2, V2, F1, g2, stateEd, v1d, W0, kb, invE;
  DevBuf<double2> h0R, E3, U2, U1d, phE, phP;
  DevBuf<double> A, Bd;   // FP64 phase rows e^{iq3.Rp}, cached panels
  DevBuf<float> Af, Bf;   // FP32 ones
  DevBuf<int> info, fail, rowU2off, rowNb2, rowE3, e2off, cmeta, pmeta;
  DevBuf<long long> rowOutOff;
  int rowsPad = 0;
  size_t pstride = 0, g2Stride = 0;
  long long pairs = 0;
  // kernel timing: 2 events for the batch's cacheElPh + 3 per outer point
  bool timing = false;
  std::vector<hipEvent_t> tev;
  double cacheMs = 0., coupMs = 0., redMs = 0., coupFlops = 0.;
  long long coupLaunches = 0, cacheLaunches = 0;
  ~PhelHandle() {
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
  }
};

static void phel_setup(PhelHandle* P, const phx_problem& p, ElphHandle* H, const int* qmesh) {
  PHX_REQUIRE(H, "phx_phel_setup: null coupling");
  PHX_REQUIRE(p.nw == H->nw && p.nm == H->nm, "phx_phel_setup: coupling and band structures differ in size");
  PHX_REQUIRE(p.nCalc >= 1 && p.nCalc <= kCMax, "phx_phel_setup: 1..4 (T, mu) calculations supported");
  PHX_REQUIRE(p.smearing == PHX_SMEAR_GAUSSIAN || p.smearing == PHX_SMEAR_ADAPTIVE,
              "phx_phel_setup: unknown smearing");
  PHX_REQUIRE(p.smearing == PHX_SMEAR_ADAPTIVE || p.sigma > 0., "phx_phel_setup: sigma must be > 0");
  for (int i = 0; i < 3; ++i) PHX_REQUIRE(p.mesh[i] >= 1 && qmesh[i] >= 1, "phx_phel_setup: bad meshes");
  PHX_REQUIRE(H->nw <= 32, "phx_phel_setup: at most 32 Wannier functions");
  P->H = H;
  P->s = H->stream;
  hipStream_t s = P->s;
  const int nw = p.nw, nm = p.nm;
  P->nw = nw;
  P->nm = nm;
  P->nCalc = p.nCalc;
  P->nk = p.mesh[0] * p.mesh[1] * p.mesh[2];
  P->nq = qmesh[0] * qmesh[1] * qmesh[2];
  for (int i = 0; i < 3; ++i) {
    P->qmesh[i] = qmesh[i];
    P->kmesh[i] = p.mesh[i];
  }
  P->T.assign(p.temperatures, p.temperatures + p.nCalc);
  P->mu.assign(p.chemPots, p.chemPots + p.nCalc);
  // norm = 1 / prod(kMesh) (phel_scattering.cpp:61)
  P->norm = 1. / (double(p.mesh[0]) * p.mesh[1] * p.mesh[2]);
  P->adaptive = p.smearing == PHX_SMEAR_ADAPTIVE;
  if (!P->adaptive) {
    P->invSigma = 1. / p.sigma;
    P->deltaPref = 1. / p.sigma / std::sqrt(kPi);
  }
  // k2 = k1 + q3 lies on the k mesh when the q mesh divides it: H(k2) = H(k2 - G)
  // for the Wannier Hamiltonian, so the k-mesh eigenpairs serve every k2 (the
  // phases of the two Fourier sums agree to rounding).  Adaptive smearing keeps
  // the unfolded points, whose Gamma rule depends on |k2|.
  P->folded = !P->adaptive && p.mesh[0] % qmesh[0] == 0 && p.mesh[1] % qmesh[1] == 0 &&
              p.mesh[2] % qmesh[2] == 0;
  // ---- electrons on the k mesh and the window
  P->kcart.resize(3 * size_t(P->nk));
  for (int ik = 0; ik < P->nk; ++ik) mesh_cart(p.recip, p.mesh, ik, &P->kcart[3 * size_t(ik)]);
  h2d(P->Rh, p.Rh, size_t(p.nRh) * 3, s);
  h2d(P->degh, p.degh, size_t(p.nRh), s);
  h2d(P->h0R, reinterpret_cast<const double2*>(p.h0R), size_t(p.nRh) * nw * nw, s);
  std::vector<double> Eall(size_t(P->nk) * nw), Uall(2 * size_t(P->nk) * nw * nw);
  {
    DevBuf<double> dk;
    h2d(dk, P->kcart, s);
    DevBuf<double2> U(size_t(P->nk) * nw * nw);
    DevBuf<double> E(size_t(P->nk) * nw);
    DevBuf<int> inf(P->nk);
    launch_bloch(false, P->nk, nw, p.nRh, dk.p, P->Rh.p, P->degh.p, P->h0R.p, U.p, s);
    launch_jacobi(nw, P->nk, U.p, E.p, inf.p, nullptr, s);
    std::vector<int> hi(P->nk);
    PHX_HIP(hipMemcpyAsync(Eall.data(), E.p, E.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipMemcpyAsync(Uall.data(), U.p, U.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipMemcpyAsync(hi.data(), inf.p, inf.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < P->nk; ++i) PHX_REQUIRE(hi[i] >= 0, "electron eigensolve did not converge");
    if (P->folded) {  // the k-mesh eigenpairs become the k2 pools
      P->U2 = std::move(U);
      P->E2 = std::move(E);
    }
  }
  for (int ik = 0; ik < P->nk; ++ik) {
    int l = -1, h = -1;
    for (int ib = 0; ib < nw; ++ib) {
      const double e = Eall[size_t(ik) * nw + ib];
      if (e < p.winMax && e > p.winMin) {
        if (l < 0) l = ib;
        h = ib;
      }
    }
    if (l < 0) continue;
    P->active.push_back(ik);
    P->lo.push_back(l);
    P->nb.push_back(h - l + 1);
  }
  P->na = int(P->active.size());
  PHX_REQUIRE(P->na > 0, "energy window selects no states");
  for (int a = 0; a < P->na; ++a) {
    const int ik = P->active[a], n = P->nb[a];
    P->stateOff.push_back(P->stateE.size());
    P->u1Off.push_back(P->U1.size() / 2);
    for (int b = 0; b < n; ++b) {
      P->stateE.push_back(Eall[size_t(ik) * nw + P->lo[a] + b]);
      const double* col = &Uall[2 * (size_t(ik) * nw * nw + size_t(nw) * (P->lo[a] + b))];
      P->U1.insert(P->U1.end(), col, col + 2 * nw);
    }
    P->nb1max = std::max(P->nb1max, n);
  }
  h2d(P->U1d, reinterpret_cast<const double2*>(P->U1.data()), P->U1.size() / 2, s);
  h2d(P->stateEd, P->stateE, s);
  if (P->adaptive) {
    // k1 velocities as the electron band structure holds them (single-point
    // velocity operator, active_bandstructure.cpp:1032-1071)
    std::vector<double> kc(3 * size_t(P->na));
    for (int a = 0; a < P->na; ++a)
      for (int i = 0; i < 3; ++i) kc[3 * size_t(a) + i] = P->kcart[3 * size_t(P->active[a]) + i];
    std::vector<double> E(size_t(P->na) * nw), U(2 * size_t(P->na) * nw * nw),
        V(2 * size_t(P->na) * nw * nw * 3);
    diag_velocities(false, P->na, kc.data(), nw, p.nRh, p.Rh, p.degh, p.h0R, nullptr, nullptr,
                    PHX_GAMMA_POINT, E.data(), U.data(), V.data());
    for (int a = 0; a < P->na; ++a)
      for (int b = 0; b < P->nb[a]; ++b) {
        const int bf = P->lo[a] + b;
        for (int i = 0; i < 3; ++i)
          P->v1.push_back(V[2 * ((((size_t(a) * nw + bf) * nw) + bf) * 3 + i)]);
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) P->qT[3 * i + j] = p.recip[i + 3 * j] / p.mesh[i];
    h2d(P->v1d, P->v1, s);
    P->hRh.assign(p.Rh, p.Rh + size_t(p.nRh) * 3);
    P->hdegh.assign(p.degh, p.degh + p.nRh);
    P->hh0R.assign(p.h0R, p.h0R + 2 * size_t(p.nRh) * nw * nw);
    P->V2.alloc(size_t(P->nq) * nw * 3);
  }
  // ---- phonons on the q mesh (PhononH0 with the long-range term)
  P->qcart.resize(3 * size_t(P->nq));
  for (int iq = 0; iq < P->nq; ++iq) mesh_cart(p.recip, qmesh, iq, &P->qcart[3 * size_t(iq)]);
  {
    DevBuf<double> dq, dRf, dwf, dmat, dmass;
    h2d(dq, P->qcart, s);
    h2d(dRf, p.Rf, size_t(p.nRf) * 3, s);
    h2d(dwf, p.wf, size_t(p.nRf), s);
    h2d(dmat, p.mat2R, size_t(p.nRf) * nm * nm, s);
    h2d(dmass, p.masses, size_t(nm), s);
    P->E3.alloc(size_t(P->nq) * nm * nm);
    P->W3.alloc(size_t(P->nq) * nm);
    DevBuf<int> qinfo(P->nq);
    launch_bloch(true, P->nq, nm, p.nRf, dq.p, dRf.p, dwf.p, dmat.p, P->E3.p, s);
    DevLongRange dlr;
    if (p.nG > 0) {
      PHX_REQUIRE(3 * p.natoms == nm, "phx_problem: natoms * 3 != nm with a long-range term");
      dlr.upload(p.natoms, p.nG, p.atomPos, p.bornCharges, p.dielectric, p.gVectors, p.lrConstant,
                 p.lrNorm, p.gMax, s);
      launch_ph_long_range(P->nq, dq.p, dlr.lr, P->E3.p, s);
    }
    launch_jacobi(nm, P->nq, P->E3.p, P->W3.p, qinfo.p, dmass.p, s);
    std::vector<int> qi(P->nq);
    P->omega.resize(size_t(P->nq) * nm);
    PHX_HIP(hipMemcpyAsync(qi.data(), qinfo.p, qinfo.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipMemcpyAsync(P->omega.data(), P->W3.p, P->W3.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < P->nq; ++i) PHX_REQUIRE(qi[i] >= 0, "phonon eigensolve did not converge");
    // phase rows e^{iq3.Rp} (explicit-q mode of the coupling kernel)
    const int rows = coupling_rows_per_tile_f64(H->ldb);
    P->rowsPad = ceil_div(P->nq, rows) * rows;
    if (H->precision == PHX_FP64) {
      P->A.alloc(size_t(P->rowsPad) * H->Kp);
      launch_phase_rows<double>(P->nq, P->rowsPad, H->nRp, H->Kp, dq.p, H->Rp.p, P->A.p, s);
    } else {
      P->Af.alloc(size_t(P->rowsPad) * H->Kp);
      launch_phase_rows<float>(P->nq, P->rowsPad, H->nRp, H->Kp, dq.p, H->Rp.p, P->Af.p, s);
    }
  }
  // ---- batched cacheElPh: e^{ik1.Re}/de per outer point, 1/dp (the e^{iq3.Rp}
  // factor sits in the phase rows), panels of up to kCacheNKK outer points
  {
    std::vector<double> degE(H->nRe), degP(H->nRp);
    PHX_HIP(hipMemcpyAsync(degE.data(), H->degE.p, H->degE.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipMemcpyAsync(degP.data(), H->degP.p, H->degP.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    std::vector<double> inv(H->nRe);
    for (int i = 0; i < H->nRe; ++i) inv[i] = 1. / degE[i];
    h2d(P->invE, inv, s);
    std::vector<double2> pp(size_t(kCacheNKK) * H->nRp);
    for (int k = 0; k < kCacheNKK; ++k)
      for (int i = 0; i < H->nRp; ++i) pp[size_t(k) * H->nRp + i] = make_double2(1. / degP[i], 0.);
    h2d(P->phP, pp, s);
    P->phE.alloc(size_t(kCacheNKK) * H->nRe);
    P->kb.alloc(3 * size_t(kCacheNKK));
    P->cmeta.alloc(3 * size_t(kCacheNKK));
    P->pmeta.alloc(size_t(kCacheNKK) * (1 + 2 * P->nb1max));
    P->pstride = size_t(H->Kp) * H->ldb;
    const size_t nPanels = size_t(kCacheNKK) * P->nb1max;
    if (H->precision == PHX_FP64) {
      P->Bd.alloc(P->pstride * nPanels);
      launch_fill<double>(P->Bd.p, P->Bd.n, 0., s);  // padding stays zero
    } else {
      P->Bf.alloc(P->pstride * nPanels);
      launch_fill<float>(P->Bf.p, P->Bf.n, 0.f, s);
    }
  }
  // ---- coupling rows (one per q3) and the reduction's buffers
  P->g2Stride = size_t(P->nb1max) * nw * nm;
  std::vector<int> u2off(P->nq), nb2(P->nq, nw), e3i(P->nq), e2o(P->nq);
  std::vector<long long> oo(P->nq);
  for (int q = 0; q < P->nq; ++q) {
    u2off[q] = q * nw * nw;  // unfolded k2: per outer point eigenpairs of nq points
    e2o[q] = q * nw;
    e3i[q] = q;
    oo[q] = (long long)q * (long long)P->g2Stride;
  }
  h2d(P->rowU2off, u2off, s);
  h2d(P->e2off, e2o, s);
  h2d(P->rowNb2, nb2, s);
  h2d(P->rowE3, e3i, s);
  h2d(P->rowOutOff, oo, s);
  if (!P->folded) {
    P->U2.alloc(size_t(P->nq) * nw * nw);
    P->E2.alloc(size_t(P->nq) * nw);
    P->k2.alloc(size_t(P->nq) * 3);
    P->info.alloc(P->nq);
  }
  P->g2.alloc(size_t(P->nq) * P->g2Stride);
  P->F1.alloc(size_t(kCacheNKK) * kCMax * nw);
  P->W0.alloc(size_t(nm));
  launch_fill<double>(P->W0.p, P->W0.n, 0., s);
  P->fail.alloc(1);
  launch_fill<int>(P->fail.p, 1, 0, s);
  std::vector<double> it(P->nCalc);
  for (int c = 0; c < P->nCalc; ++c) it[c] = 1. / P->T[c];
  h2d(P->invT, it, s);
  P->lw.alloc(size_t(P->nCalc) * P->nq * nm);
  launch_fill<double>(P->lw.p, P->lw.n, 0., s);
  PHX_HIP(hipStreamSynchronize(s));
}

// Up to kCacheNKK outer points (active indices as[0..n)): one batched
// cacheElPh (one gW stream for all of them), then per outer point the
// coupling of every q3 in one launch and the symmetrize + rate kernels.
// Enqueued on the handle's stream; host-side buffers of pageable memory are
// staged by the copies, so nothing waits for the device here.
static void phel_batch(PhelHandle* P, const int* as, int n) {
  ElphHandle* H = P->H;
  hipStream_t s = P->s;
  const int nw = P->nw, nm = P->nm, nq = P->nq, K = kCacheNKK, blk = 1 + 2 * P->nb1max;
  const bool f64 = H->precision == PHX_FP64;
  if (P->timing && P->tev.empty()) {
    P->tev.resize(2 + 3 * size_t(K));
    for (hipEvent_t& e : P->tev) PHX_HIP(hipEventCreate(&e));
  }
  auto rec = [&](int j) {
    if (P->timing) PHX_HIP(hipEventRecord(P->tev[j], s));
  };
  // ---- batched cacheElPh
  std::vector<double> kb(3 * size_t(n));
  std::vector<int> cm(3 * size_t(K), 0), pm(size_t(K) * blk, 0);
  std::vector<double> F(size_t(K) * kCMax * nw, 0.);
  int pb = 0;
  for (int i = 0; i < n; ++i) {
    const int a = as[i], nb1 = P->nb[a];
    for (int d = 0; d < 3; ++d) kb[3 * size_t(i) + d] = P->kcart[3 * size_t(P->active[a]) + d];
    cm[i] = int(P->u1Off[a]);
    cm[K + i] = nb1;
    cm[2 * K + i] = pb;
    pb += nb1;
    int* m = &pm[size_t(i) * blk];  // kkNb1, panelKK (all 0), panelB1
    m[0] = nb1;
    for (int b = 0; b < nb1; ++b) m[1 + P->nb1max + b] = b;
    const double* e1 = &P->stateE[P->stateOff[a]];
    for (int c = 0; c < P->nCalc; ++c)
      for (int b = 0; b < nb1; ++b) {
        // Particle::getPopPopPm1 (particle.cpp:117-128), fermions
        const double x = 1. / std::cosh((e1[b] - P->mu[c]) / 2. / P->T[c]);
        F[size_t(i) * kCMax * nw + size_t(c) * nb1 + b] = 0.25 * x * x;  // (calc, b1) of point i
      }
  }
  PHX_HIP(hipMemcpyAsync(P->kb.p, kb.data(), kb.size() * sizeof(double), hipMemcpyHostToDevice, s));
  PHX_HIP(hipMemcpyAsync(P->cmeta.p, cm.data(), cm.size() * sizeof(int), hipMemcpyHostToDevice, s));
  PHX_HIP(hipMemcpyAsync(P->pmeta.p, pm.data(), pm.size() * sizeof(int), hipMemcpyHostToDevice, s));
  PHX_HIP(hipMemcpyAsync(P->F1.p, F.data(), F.size() * sizeof(double), hipMemcpyHostToDevice, s));
  rec(0);
  launch_phase_table(n, H->nRe, P->kb.p, H->Re.p, P->invE.p, 1.0, P->phE.p, s);
  CacheParams cp{};
  cp.nRe = H->nRe;
  cp.nRp = H->nRp;
  cp.nm = nm;
  cp.nw = nw;
  cp.nkk = n;
  cp.gW = H->gw_ptr();
  cp.phE = P->phE.p;
  cp.phP = P->phP.p;
  cp.U1pool = P->U1d.p;
  cp.U1off = P->cmeta.p;
  cp.nb1 = P->cmeta.p + K;
  cp.panelBase = P->cmeta.p + 2 * K;
  cp.panelStride = P->pstride;
  cp.ldb = H->ldb;
  if (f64) {
    cp.B = P->Bd.p;
    launch_cache<double, double>(cp, s);
  } else {
    cp.B = P->Bf.p;
    launch_cache<float, float>(cp, s);
  }
  rec(1);
  // ---- every outer point against all q3
  std::vector<double> k2(3 * size_t(nq));
  std::vector<int> u2off, e2o;
  for (int i = 0; i < n; ++i) {
    const int a = as[i], nb1 = P->nb[a], ik = P->active[a];
    const double* k1 = &P->kcart[3 * size_t(ik)];
    if (P->folded) {
      // k2 = k1 + q3 on the k mesh (first index fastest)
      u2off.resize(nq);
      e2o.resize(nq);
      const int* m = P->kmesh;
      const int j0 = ik % m[0], j1 = (ik / m[0]) % m[1], j2 = ik / (m[0] * m[1]);
      for (int q = 0; q < nq; ++q) {
        const int* mq = P->qmesh;
        const int q0 = q % mq[0], q1 = (q / mq[0]) % mq[1], q2 = q / (mq[0] * mq[1]);
        const int x = (j0 + q0 * (m[0] / mq[0])) % m[0], y = (j1 + q1 * (m[1] / mq[1])) % m[1],
                  z = (j2 + q2 * (m[2] / mq[2])) % m[2];
        const int k = (z * m[1] + y) * m[0] + x;
        u2off[q] = k * nw * nw;
        e2o[q] = k * nw;
      }
      PHX_HIP(hipMemcpyAsync(P->rowU2off.p, u2off.data(), nq * sizeof(int), hipMemcpyHostToDevice, s));
      PHX_HIP(hipMemcpyAsync(P->e2off.p, e2o.data(), nq * sizeof(int), hipMemcpyHostToDevice, s));
    } else {
      // k2 = k1 + q3, not folded, all bands (electronH0.populate, phel_scattering.cpp:225-243)
      for (int q = 0; q < nq; ++q)
        for (int d = 0; d < 3; ++d) k2[3 * size_t(q) + d] = P->qcart[3 * size_t(q) + d] + k1[d];
      if (P->adaptive) {
        // k2 velocities of the batched form (kokkosPopulate with velocities)
        std::vector<double> E(size_t(nq) * nw), U(2 * size_t(nq) * nw * nw), V(2 * size_t(nq) * nw * nw * 3);
        diag_velocities(false, nq, k2.data(), nw, int(P->hdegh.size()), P->hRh.data(), P->hdegh.data(),
                        P->hh0R.data(), nullptr, nullptr, PHX_GAMMA_BATCHED, E.data(), U.data(), V.data());
        std::vector<double> gv(size_t(nq) * nw * 3);
        for (int q = 0; q < nq; ++q)
          for (int b = 0; b < nw; ++b)
            for (int d = 0; d < 3; ++d)
              gv[(size_t(q) * nw + b) * 3 + d] = V[2 * ((((size_t(q) * nw + b) * nw) + b) * 3 + d)];
        PHX_HIP(hipMemcpyAsync(P->V2.p, gv.data(), gv.size() * sizeof(double), hipMemcpyHostToDevice, s));
        PHX_HIP(hipMemcpyAsync(P->E2.p, E.data(), E.size() * sizeof(double), hipMemcpyHostToDevice, s));
        PHX_HIP(hipMemcpyAsync(P->U2.p, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice, s));
      } else {
        PHX_HIP(hipMemcpyAsync(P->k2.p, k2.data(), k2.size() * sizeof(double), hipMemcpyHostToDevice, s));
        launch_bloch(false, nq, nw, int(P->degh.n), P->k2.p, P->Rh.p, P->degh.p, P->h0R.p, P->U2.p, s);
        launch_jacobi(nw, nq, P->U2.p, P->E2.p, P->info.p, nullptr, s);
        launch_phel_flag_failures(nq, P->info.p, P->fail.p, s);
      }
    }
    // calcCouplingSquared, store mode, every q3 in one launch
    const int* pmk = P->pmeta.p + size_t(i) * blk;
    CoupParams c{};
    c.lda = H->Kp;
    c.Kp = H->Kp;
    c.ldb = H->ldb;
    c.panelStride = P->pstride;
    c.M = nq;
    c.nm = nm;
    c.nw = nw;
    c.nPanels = nb1;
    c.nMtiles = P->rowsPad / coupling_rows_per_tile_f64(H->ldb);
    c.panelKK = pmk + 1;
    c.panelB1 = pmk + 1 + P->nb1max;
    c.kkNb1 = pmk;
    c.U2pool = P->U2.p;
    c.rowU2off = P->rowU2off.p;
    c.rowNb2 = P->rowNb2.p;
    c.E3pool = P->E3.p;
    c.Wpool = P->W0.p;
    c.qmode = 0;
    c.rowE3 = P->rowE3.p;
    c.out = P->g2.p;
    c.rowOutOff = P->rowOutOff.p;
    c.nb2max = nw;
    const size_t panel0 = size_t(cm[2 * K + i]) * P->pstride;
    rec(2 + 3 * i);
    if (f64) {
      c.A = P->A.p;
      c.B = P->Bd.p + panel0;
      launch_coupling<double>(c, false, s);
    } else {
      c.A = P->Af.p;
      c.B = P->Bf.p + panel0;
      launch_coupling<float>(c, false, s);
    }
    rec(3 + 3 * i);
    PhelParams r{};
    r.nq = nq;
    r.nb1 = nb1;
    r.nb2 = nw;
    r.nm = nm;
    r.nCalc = P->nCalc;
    r.g2 = P->g2.p;
    r.g2Stride = P->g2Stride;
    r.e1 = P->stateEd.p + P->stateOff[a];
    r.E2 = P->E2.p;
    r.e2off = P->e2off.p;
    r.W3 = P->W3.p;
    r.F1 = P->F1.p + size_t(i) * kCMax * nw;
    r.invT = P->invT.p;
    r.norm = P->norm;
    r.cutoff = phonon_cutoff();
    r.invSigma = P->invSigma;
    r.deltaPref = P->deltaPref;
    r.adaptive = P->adaptive;
    if (P->adaptive) {
      r.v1 = P->v1d.p + 3 * P->stateOff[a];
      r.V2 = P->V2.p;
      for (int d = 0; d < 9; ++d) r.qT[d] = P->qT[d];
      r.minWidth = 0.0001 / units::energyRyToEv;
    }
    r.lw = P->lw.p;
    launch_phel_reduce(r, s);
    rec(4 + 3 * i);
    P->pairs += nq;
  }
  if (P->timing) {
    PHX_HIP(hipEventSynchronize(P->tev[1 + 3 * size_t(n)]));
    auto ms = [&](int a, int b) {
      float t = 0.f;
      PHX_HIP(hipEventElapsedTime(&t, P->tev[a], P->tev[b]));
      return double(t);
    };
    P->cacheMs += ms(0, 1);
    P->cacheLaunches += 1;
    for (int i = 0; i < n; ++i) {
      P->coupMs += ms(2 + 3 * i, 3 + 3 * i);
      P->redMs += ms(3 + 3 * i, 4 + 3 * i);
      // R_ph transform of the coupling: real GEMM nq x 2 nRp x 2 nm nw per panel
      P->coupFlops += 8. * nq * H->nRp * nm * nw * P->nb[as[i]];
    }
    P->coupLaunches += n;
  }
}

// Outer points in batches of kCacheNKK; the eigensolver status of unfolded k2
// points is checked once at the end.
static void phel_run(PhelHandle* P, int nk1, const int* k1_list) {
  std::vector<int> list;
  if (nk1 < 0) {
    list.resize(P->na);
    std::iota(list.begin(), list.end(), 0);
  } else {
    PHX_REQUIRE(k1_list || nk1 == 0, "phx_phel_run: null k1 list");
    for (int i = 0; i < nk1; ++i)
      PHX_REQUIRE(k1_list[i] >= 0 && k1_list[i] < P->na, "phx_phel_run: k1 index out of range");
    list.assign(k1_list, k1_list + nk1);
  }
  for (size_t i = 0; i < list.size(); i += kCacheNKK)
    phel_batch(P, list.data() + i, int(std::min<size_t>(kCacheNKK, list.size() - i)));
  int bad = 0;
  PHX_HIP(hipMemcpyAsync(&bad, P->fail.p, sizeof(int), hipMemcpyDeviceToHost, P->s));
  PHX_HIP(hipStreamSynchronize(P->s));
  PHX_REQUIRE(bad == 0, "electron eigensolve did not converge");
}

static void phel_get(PhelHandle* P, double* out, int average) {
  const size_t ns = size_t(P->nq) * P->nm;
  std::vector<double> lw(size_t(P->nCalc) * ns);
  PHX_HIP(hipStreamSynchronize(P->s));
  PHX_HIP(hipMemcpy(lw.data(), P->lw.p, lw.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (average) {
    // degeneracyAveragingLinewidths over the phonon band structure, row 0
    // (scattering.cpp:1168-1207)
    for (int q = 0; q < P->nq; ++q) {
      const double* w = &P->omega[size_t(q) * P->nm];
      std::vector<double> tmp(P->nm);
      for (int b1 = 0; b1 < P->nm; ++b1) {
        double t2 = 0.;
        int n = 0;
        for (int b2 = 0; b2 < P->nm; ++b2)
          if (std::abs(w[b2] - w[b1]) < 1.0e-6) {
            n++;
            t2 = t2 + lw[size_t(q) * P->nm + b2];
          }
        tmp[b1] = t2 / float(n);
      }
      for (int b1 = 0; b1 < P->nm; ++b1) lw[size_t(q) * P->nm + b1] = tmp[b1];
    }
  }
  std::copy(lw.begin(), lw.end(), out);
}

static void lw_get(LwHandle* L, double* out, int average) {
  const int st = L->numStates;
  std::vector<double> lw(size_t(L->nCalc) * st);
  PHX_HIP(hipStreamSynchronize(L->stream));
  PHX_HIP(hipMemcpy(lw.data(), L->lw.p, lw.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (average) {
    // ScatteringMatrix::degeneracyAveragingLinewidths (scattering.cpp:1168-1207), row 0
    for (int a = 0; a < L->na; ++a) {
      const int off = L->stateOff[a], nb = L->nb[a];
      std::vector<double> tmp(nb);
      for (int b1 = 0; b1 < nb; ++b1) {
        double ekk = L->stateE[off + b1], t2 = 0.;
        int n = 0;
        for (int b2 = 0; b2 < nb; ++b2)
          if (std::abs(L->stateE[off + b2] - ekk) < 1.0e-6) {
            n++;
            t2 = t2 + lw[off + b2];
          }
        tmp[b1] = t2 / float(n);
      }
      for (int b1 = 0; b1 < nb; ++b1) lw[off + b1] = tmp[b1];
    }
  }
  std::copy(lw.begin(), lw.end(), out);
}

}  // namespace phx

using namespace phx;

// ------------------------------------------------------------ fixed device scratch
// Scratch of phx_el_bloch_diag_noalloc: one fixed array on every GPU the
// library runs on, held until the process ends, so that the entry point
// allocates nothing.  In host code `scratch` names a host-side stand-in, not
// device memory: kernels, copies and fills use the device address that
// hipGetSymbolAddress returns, looked up once per device on the entry
// point's first call there (each GPU has its own copy at its own address).
// Every call on a GPU uses the same array: calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex gScratchMutex;
unsigned char* gScratchDev[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device.
unsigned char* scratch_on_current_device() {
  int dev = 0;
  PHX_HIP(hipGetDevice(&dev));
  PHX_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "fixed scratch: device index out of range");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratchDev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    PHX_REQUIRE(e == hipSuccess && p, std::string("fixed scratch: hipGetSymbolAddress failed: ") +
                                          hipGetErrorString(e));
    gScratchDev[dev] = static_cast<unsigned char*>(p);
  }
  return gScratchDev[dev];
}
}  // namespace

// ============================================================ C ABI
struct phx_elph : phx::ElphHandle {};
struct phx_lw : phx::LwHandle {};
struct phx_elph_file : phx::ElphFile {
  using phx::ElphFile::ElphFile;
};
struct phx_ph_model {
  phx::PhononModel m;
};
struct phx_phel : phx::PhelHandle {};
struct phx_el_model {
  phx::ElectronModel m;
};

extern "C" {

const char* phx_last_error(void) { return phx::last_error(); }
const char* phx_version(void) { return "phx_elph 0.1 (gfx950)"; }

int phx_device_count(int* count) {
  return guard([&] {
    PHX_REQUIRE(count, "phx_device_count: null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int phx_zheev_batched(int M, int N, double* A, double* W, int* info) {
  return guard([&] {
    PHX_REQUIRE(M >= 0 && N >= 1 && N <= 44, "phx_zheev_batched: 1 <= N <= 44");
    if (M) zheev_batched(M, N, A, W, info);
  });
}

int phx_zheev_batched_dev(int M, int N, double* A_dev, double* W_dev, int* info_dev, void* stream) {
  return guard([&] {
    PHX_REQUIRE(M >= 0 && N >= 1 && N <= 44, "phx_zheev_batched_dev: 1 <= N <= 44");
    launch_jacobi(N, M, reinterpret_cast<double2*>(A_dev), W_dev, info_dev, nullptr,
                  static_cast<hipStream_t>(stream));
  });
}

int phx_el_bloch_diag(int nk, const double* k, int nw, int nR, const double* R, const double* deg,
                      const double* h0R, double* energies, double* eigvecs) {
  return guard([&] {
    PHX_REQUIRE(nk >= 0 && nw >= 1 && nw <= 44 && nR >= 1, "phx_el_bloch_diag: bad sizes");
    if (nk) bloch_diag(false, nk, k, nw, nR, R, deg, h0R, nullptr, 0, energies, eigvecs);
  });
}

int phx_el_bloch_diag_noalloc(int nk, const double* k, int nw, int nR, const double* R,
                              const double* deg, const double* h0R, double* energies, double* eigvecs) {
  return guard([&] {
    PHX_REQUIRE(nk >= 0 && nw >= 1 && nw <= 44 && nR >= 1, "phx_el_bloch_diag_noalloc: bad sizes");
    // the same device arrays as phx_el_bloch_diag, placed in the fixed
    // scratch at 256-byte boundaries
    size_t used = 0;
    auto take = [&](size_t bytes) {
      const size_t at = used;
      used += (bytes + 255) / 256 * 256;
      return at;
    };
    const size_t nE = size_t(nw) * nw;
    const size_t oK = take(size_t(nk) * 3 * sizeof(double)), oR = take(size_t(nR) * 3 * sizeof(double)),
                 oDeg = take(size_t(nR) * sizeof(double)), oCoef = take(size_t(nR) * nE * sizeof(double2)),
                 oH = take(size_t(nk) * nE * sizeof(double2)), oW = take(size_t(nk) * nw * sizeof(double)),
                 oInfo = take(size_t(nk) * sizeof(int));
    PHX_REQUIRE(used <= kNoAllocScratchBytes,
                "phx_el_bloch_diag_noalloc: the call needs " + std::to_string(used) +
                    " bytes of device scratch, more than its fixed " + std::to_string(kNoAllocScratchBytes) +
                    " (use phx_el_bloch_diag, or fewer points per call)");
    if (!nk) return;
    PHX_REQUIRE(k && R && deg && h0R && energies && eigvecs, "phx_el_bloch_diag_noalloc: null argument");
    unsigned char* base = scratch_on_current_device();
    auto* dk = reinterpret_cast<double*>(base + oK);
    auto* dR = reinterpret_cast<double*>(base + oR);
    auto* dDeg = reinterpret_cast<double*>(base + oDeg);
    auto* dCoef = reinterpret_cast<double2*>(base + oCoef);
    auto* dH = reinterpret_cast<double2*>(base + oH);
    auto* dW = reinterpret_cast<double*>(base + oW);
    auto* dInfo = reinterpret_cast<int*>(base + oInfo);
    const hipStream_t s = 0;
    PHX_HIP(hipMemcpyAsync(dk, k, size_t(nk) * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    PHX_HIP(hipMemcpyAsync(dR, R, size_t(nR) * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    PHX_HIP(hipMemcpyAsync(dDeg, deg, size_t(nR) * sizeof(double), hipMemcpyHostToDevice, s));
    PHX_HIP(hipMemcpyAsync(dCoef, h0R, size_t(nR) * nE * sizeof(double2), hipMemcpyHostToDevice, s));
    // k_bloch writes every entry of H and k_jacobi every eigenvalue and every
    // info flag, so nothing left from an earlier call is read
    launch_bloch(false, nk, nw, nR, dk, dR, dDeg, dCoef, dH, s);
    launch_jacobi(nw, nk, dH, dW, dInfo, nullptr, s);
    PHX_HIP(hipMemcpy(energies, dW, size_t(nk) * nw * sizeof(double), hipMemcpyDeviceToHost));
    PHX_HIP(hipMemcpy(eigvecs, dH, size_t(nk) * nE * sizeof(double2), hipMemcpyDeviceToHost));
    // convergence flags through a stack buffer (no host allocation either)
    int inf[1024];
    for (int i0 = 0; i0 < nk; i0 += 1024) {
      const int n = std::min(1024, nk - i0);
      PHX_HIP(hipMemcpy(inf, dInfo + i0, size_t(n) * sizeof(int), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i)
        if (inf[i] < 0) throw Error(-4, "Bloch eigensolve did not converge");
    }
  });
}

int phx_ph_bloch_diag(int nq, const double* q, int nm, int nR, const double* R, const double* w,
                      const double* mat2R, const double* masses, int mass_scaling, double* omega,
                      double* eigvecs) {
  return guard([&] {
    PHX_REQUIRE(nq >= 0 && nm >= 1 && nm <= 44 && nR >= 1, "phx_ph_bloch_diag: bad sizes");
    if (nq) bloch_diag(true, nq, q, nm, nR, R, w, mat2R, masses, mass_scaling, omega, eigvecs);
  });
}

int phx_bloch_diag_timed(int phonon, int nk, const double* k, int n, int nR, const double* R, const double* w,
                         const double* coef, const double* masses, int reps, double* ms, double* energies,
                         double* eigvecs) {
  return guard([&] {
    PHX_REQUIRE(nk >= 1 && n >= 1 && n <= 44 && nR >= 1 && reps >= 1 && ms && (!phonon || masses),
                "phx_bloch_diag_timed: bad arguments");
    bloch_diag(phonon != 0, nk, k, n, nR, R, w, coef, masses, 1, energies, eigvecs, nullptr, reps, ms);
  });
}

int phx_elph_create(phx_elph** h, const double* gW, int nw, int nm, int nRp, int nRe,
                    const double* elR, const double* elDeg, const double* phR, const double* phDeg,
                    int precision) {
  return guard([&] {
    PHX_REQUIRE(h && gW, "phx_elph_create: null argument");
    auto* H = new phx_elph();
    try {
      elph_common_init(H, nw, nm, nRp, nRe, elR, elDeg, phR, phDeg, precision);
      const size_t n = size_t(nRe) * nRp * nm * nw * nw;
      h2d(H->gW, reinterpret_cast<const double2*>(gW), n, H->stream);
      if (precision != PHX_FP64) {
        H->gWf.alloc(n);
        launch_to_float(reinterpret_cast<const double*>(H->gW.p), reinterpret_cast<float*>(H->gWf.p),
                        2 * n, H->stream);
        PHX_HIP(hipStreamSynchronize(H->stream));
        H->gW.release();
      }
      PHX_HIP(hipStreamSynchronize(H->stream));
    } catch (...) {
      delete H;
      throw;
    }
    *h = H;
  });
}

int phx_elph_create_synthetic(phx_elph** h, int nw, int nm, int nRp, int nRe, const double* elR,
                              const double* elDeg, const double* phR, const double* phDeg,
                              const double* decay, unsigned long long seed, int precision) {
  return guard([&] {
    PHX_REQUIRE(h && decay, "phx_elph_create_synthetic: null argument");
    auto* H = new phx_elph();
    try {
      elph_common_init(H, nw, nm, nRp, nRe, elR, elDeg, phR, phDeg, precision);
      DevBuf<double> dec;
      h2d(dec, decay, size_t(nRe) * nRp, H->stream);
      const size_t n = size_t(nRe) * nRp * nm * nw * nw;
      if (precision == PHX_FP64) {
        H->gW.alloc(n);
        launch_synth_gw(H->gW.p, n, size_t(nm) * nw * nw, dec.p, seed, H->stream);
      } else {
        H->gWf.alloc(n);
        launch_synth_gw_f32(H->gWf.p, n, size_t(nm) * nw * nw, dec.p, seed, H->stream);
      }
      PHX_HIP(hipStreamSynchronize(H->stream));
    } catch (...) {
      delete H;
      throw;
    }
    *h = H;
  });
}

int phx_ph_model_read_qe(phx_ph_model** m, const char* fc_path, const char* sum_rule) {
  return guard([&] {
    PHX_REQUIRE(m && fc_path, "phx_ph_model_read_qe: null argument");
    *m = nullptr;
    auto* M = new phx_ph_model();
    try {
      M->m = phx::read_qe_fc(fc_path, sum_rule ? sum_rule : "");
      PHX_REQUIRE(M->m.nm <= 36, "phx_ph_model_read_qe: at most 12 atoms");
    } catch (...) {
      delete M;
      throw;
    }
    *m = M;
  });
}

int phx_ph_model_dims(phx_ph_model* m, int* d) {
  return guard([&] {
    PHX_REQUIRE(m && d, "phx_ph_model_dims: null argument");
    const phx::PhononModel& x = m->m;
    const int v[8] = {x.nat, x.nm, x.nR, x.nG, x.hasDielectric ? 1 : 0, x.qGrid[0], x.qGrid[1], x.qGrid[2]};
    std::memcpy(d, v, sizeof(v));
  });
}

int phx_ph_model_crystal(phx_ph_model* m, double* cell9, double* recip9, double* volume,
                         double* positions, double* masses) {
  return guard([&] {
    PHX_REQUIRE(m, "phx_ph_model_crystal: null handle");
    const phx::PhononModel& x = m->m;
    if (cell9) std::memcpy(cell9, x.cell, sizeof(x.cell));
    if (recip9) std::memcpy(recip9, x.recip, sizeof(x.recip));
    if (volume) *volume = x.volume;
    if (positions) std::copy(x.positions.begin(), x.positions.end(), positions);
    if (masses) std::copy(x.masses.begin(), x.masses.end(), masses);
  });
}

int phx_ph_model_force_constants(phx_ph_model* m, double* R, double* weights, double* mat2R) {
  return guard([&] {
    PHX_REQUIRE(m, "phx_ph_model_force_constants: null handle");
    const phx::PhononModel& x = m->m;
    if (R) std::copy(x.R.begin(), x.R.end(), R);
    if (weights) std::copy(x.weights.begin(), x.weights.end(), weights);
    if (mat2R) std::copy(x.mat2R.begin(), x.mat2R.end(), mat2R);
  });
}

int phx_ph_model_long_range(phx_ph_model* m, double* born, double* dielectric, double* gVectors,
                            double* lrConstant, double* norm_gmax2) {
  return guard([&] {
    PHX_REQUIRE(m, "phx_ph_model_long_range: null handle");
    const phx::PhononModel& x = m->m;
    if (born) std::copy(x.born.begin(), x.born.end(), born);
    if (dielectric) std::memcpy(dielectric, x.dielectric, sizeof(x.dielectric));
    if (gVectors) std::copy(x.gVectors.begin(), x.gVectors.end(), gVectors);
    if (lrConstant) std::copy(x.lrConstant.begin(), x.lrConstant.end(), lrConstant);
    if (norm_gmax2) {
      norm_gmax2[0] = x.lrNorm;
      norm_gmax2[1] = x.gMax;
    }
  });
}

int phx_ph_model_diag(phx_ph_model* m, int nq, const double* q, int mass_scaling, double* omega,
                      double* eigvecs) {
  return guard([&] {
    PHX_REQUIRE(m && (nq == 0 || (q && omega && eigvecs)), "phx_ph_model_diag: null argument");
    PHX_REQUIRE(nq >= 0, "phx_ph_model_diag: nq < 0");
    const phx::PhononModel& x = m->m;
    if (nq)
      bloch_diag(true, nq, q, x.nm, x.nR, x.R.data(), x.weights.data(), x.mat2R.data(),
                 x.masses.data(), mass_scaling, omega, eigvecs, &x);
  });
}

int phx_ph_model_diag_velocities(phx_ph_model* m, int nq, const double* q, int gamma_rule,
                                 double* omega, double* eigvecs, double* velocities) {
  return guard([&] {
    PHX_REQUIRE(m && (nq == 0 || (q && omega && eigvecs && velocities)),
                "phx_ph_model_diag_velocities: null argument");
    PHX_REQUIRE(nq >= 0, "phx_ph_model_diag_velocities: nq < 0");
    const phx::PhononModel& x = m->m;
    if (nq)
      diag_velocities(true, nq, q, x.nm, x.nR, x.R.data(), x.weights.data(), x.mat2R.data(),
                      x.masses.data(), &x, gamma_rule, omega, eigvecs, velocities);
  });
}

int phx_ph_model_destroy(phx_ph_model* m) {
  return guard([&] { delete m; });
}

int phx_el_model_read_w90(phx_el_model** m, const char* tb_path, const double* cell9) {
  return guard([&] {
    PHX_REQUIRE(m && tb_path, "phx_el_model_read_w90: null argument");
    *m = nullptr;
    auto* M = new phx_el_model();
    try {
      M->m = phx::read_w90_tb(tb_path, cell9);
      PHX_REQUIRE(M->m.nw <= 44, "phx_el_model_read_w90: at most 44 Wannier functions");
    } catch (...) {
      delete M;
      throw;
    }
    *m = M;
  });
}

int phx_el_model_dims(phx_el_model* m, int* d) {
  return guard([&] {
    PHX_REQUIRE(m && d, "phx_el_model_dims: null argument");
    d[0] = m->m.nw;
    d[1] = m->m.nR;
  });
}

int phx_el_model_arrays(phx_el_model* m, double* cell9, double* R, double* deg, double* h0R,
                        double* rMatrix) {
  return guard([&] {
    PHX_REQUIRE(m, "phx_el_model_arrays: null handle");
    const phx::ElectronModel& x = m->m;
    if (cell9) std::memcpy(cell9, x.cell, sizeof(x.cell));
    if (R) std::copy(x.R.begin(), x.R.end(), R);
    if (deg) std::copy(x.deg.begin(), x.deg.end(), deg);
    if (h0R) std::copy(x.h0R.begin(), x.h0R.end(), h0R);
    if (rMatrix) std::copy(x.rMat.begin(), x.rMat.end(), rMatrix);
  });
}

int phx_el_model_diag(phx_el_model* m, int nk, const double* k, double* energies, double* eigvecs) {
  return guard([&] {
    PHX_REQUIRE(m && (nk == 0 || (k && energies && eigvecs)), "phx_el_model_diag: null argument");
    PHX_REQUIRE(nk >= 0, "phx_el_model_diag: nk < 0");
    const phx::ElectronModel& x = m->m;
    if (nk)
      bloch_diag(false, nk, k, x.nw, x.nR, x.R.data(), x.deg.data(), x.h0R.data(), nullptr, 0,
                 energies, eigvecs);
  });
}

int phx_el_model_diag_velocities(phx_el_model* m, int nk, const double* k, int gamma_rule,
                                 double* energies, double* eigvecs, double* velocities) {
  return guard([&] {
    PHX_REQUIRE(m && (nk == 0 || (k && energies && eigvecs && velocities)),
                "phx_el_model_diag_velocities: null argument");
    PHX_REQUIRE(nk >= 0, "phx_el_model_diag_velocities: nk < 0");
    const phx::ElectronModel& x = m->m;
    if (nk)
      diag_velocities(false, nk, k, x.nw, x.nR, x.R.data(), x.deg.data(), x.h0R.data(), nullptr, nullptr,
                      gamma_rule, energies, eigvecs, velocities);
  });
}

int phx_el_model_destroy(phx_el_model* m) {
  return guard([&] { delete m; });
}

int phx_problem_set_models(phx_problem* prob, phx_el_model* el, phx_ph_model* ph) {
  return guard([&] {
    PHX_REQUIRE(prob && el && ph, "phx_problem_set_models: null argument");
    const phx::PhononModel& f = ph->m;
    const phx::ElectronModel& e = el->m;
    std::memcpy(prob->recip, f.recip, sizeof(f.recip));
    prob->nw = e.nw;
    prob->nRh = e.nR;
    prob->Rh = e.R.data();
    prob->degh = e.deg.data();
    prob->h0R = e.h0R.data();
    prob->nm = f.nm;
    prob->nRf = f.nR;
    prob->Rf = f.R.data();
    prob->wf = f.weights.data();
    prob->mat2R = f.mat2R.data();
    prob->masses = f.masses.data();
    prob->natoms = f.nat;
    prob->nG = f.hasDielectric ? f.nG : 0;
    prob->atomPos = f.positions.data();
    prob->bornCharges = f.born.data();
    prob->dielectric = f.dielectric;
    prob->gVectors = f.gVectors.data();
    prob->lrConstant = f.lrConstant.data();
    prob->lrNorm = f.lrNorm;
    prob->gMax = f.gMax;
  });
}

int phx_elph_create_from_file(phx_elph** h, const char* path, int precision) {
  return guard([&] {
    PHX_REQUIRE(h && path, "phx_elph_create_from_file: null argument");
    phx::ElphFile f(path);
    const phx::ElphHeader& hd = f.header();
    std::vector<double> eR(3 * size_t(hd.nRe)), eD(size_t(hd.nRe)), pR(3 * size_t(hd.nRp)),
        pD(size_t(hd.nRp));
    f.bravais(eR.data(), eD.data(), pR.data(), pD.data());
    auto* H = new phx_elph();
    try {
      elph_common_init(H, hd.nw, hd.nm, hd.nRp, hd.nRe, eR.data(), eD.data(), pR.data(), pD.data(),
                       precision);
      upload_gw_stream([&f](uint64_t f0, uint64_t k, double* d) { f.read_gw(f0, k, d); },
                       f.gw_elements(), precision, H->gW, H->gWf, H->stream);
    } catch (...) {
      delete H;
      throw;
    }
    *h = H;
  });
}

int phx_elph_file_open(phx_elph_file** f, const char* path) {
  return guard([&] {
    PHX_REQUIRE(f && path, "phx_elph_file_open: null argument");
    *f = nullptr;
    *f = new phx_elph_file(path);
  });
}

int phx_elph_file_header(phx_elph_file* f, int* v) {
  return guard([&] {
    PHX_REQUIRE(f && v, "phx_elph_file_header: null argument");
    const phx::ElphHeader& h = f->header();
    const int a[13] = {h.fileFormat, h.numElectrons, h.numSpin, h.nw, h.nm, h.nRe, h.nRp,
                       h.kMesh[0], h.kMesh[1], h.kMesh[2], h.qMesh[0], h.qMesh[1], h.qMesh[2]};
    std::memcpy(v, a, sizeof(a));
  });
}

int phx_elph_file_bravais(phx_elph_file* f, double* elR, double* elDeg, double* phR,
                          double* phDeg) {
  return guard([&] {
    PHX_REQUIRE(f, "phx_elph_file_bravais: null handle");
    f->bravais(elR, elDeg, phR, phDeg);
  });
}

int phx_elph_file_read_gw(phx_elph_file* f, int re_begin, int re_count, double* out) {
  return guard([&] {
    PHX_REQUIRE(f && out, "phx_elph_file_read_gw: null argument");
    const int nRe = f->header().nRe;
    PHX_REQUIRE(re_begin >= 0 && re_count >= 0 && re_begin <= nRe && re_count <= nRe - re_begin,
                "phx_elph_file_read_gw: R_e range outside the file");
    const uint64_t per = f->gw_elements_per_re();
    f->read_gw(per * uint64_t(re_begin), per * uint64_t(re_count), out);
  });
}

int phx_elph_file_close(phx_elph_file* f) {
  return guard([&] { delete f; });
}

int phx_elph_file_write(const char* path, int format, const int* v, const double* elR,
                        const double* elDeg, const double* phR, const double* phDeg,
                        const double* gW) {
  return guard([&] {
    PHX_REQUIRE(path && v, "phx_elph_file_write: null argument");
    phx::ElphHeader h;
    h.fileFormat = format;
    h.numElectrons = v[0];
    h.numSpin = v[1];
    h.nw = v[2];
    h.nm = v[3];
    h.nRe = v[4];
    h.nRp = v[5];
    for (int i = 0; i < 3; ++i) {
      h.kMesh[i] = v[6 + i];
      h.qMesh[i] = v[9 + i];
    }
    phx::write_elph_file(path, h, elR, elDeg, phR, phDeg, gW);
  });
}

int phx_elph_cache(phx_elph* h, const double* U1, int nb1, const double* k1) {
  return guard([&] {
    PHX_REQUIRE(h && U1 && k1, "phx_elph_cache: null argument");
    elph_cache(h, U1, nb1, k1);
  });
}

int phx_elph_coupling_sq(phx_elph* h, int nq, const double* U2s, const int* nb2s, const double* e3s,
                         const double* q3, const double* polar, double* out) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_elph_coupling_sq: null handle");
    PHX_REQUIRE(nq == 0 || (U2s && nb2s && e3s && q3 && out), "phx_elph_coupling_sq: null argument");
    elph_coupling_sq(h, nq, U2s, nb2s, e3s, q3, polar, out);
  });
}

int phx_polar_in_use(const double* dielectric, int num_species) {
  if (!dielectric) return 0;
  double n2 = 0.;
  for (int i = 0; i < 9; ++i) n2 += dielectric[i] * dielectric[i];
  return (n2 > 1.0e-10 && num_species > 1) ? 1 : 0;
}

int phx_elph_set_polar(phx_elph* h, const phx_polar* polar) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_elph_set_polar: null handle");
    PHX_REQUIRE(!polar || 3 * polar->natoms == h->nm, "phx_elph_set_polar: nm must be 3 natoms");
    h->polar.set(polar, h->stream);
  });
}

int phx_elph_polar_part1(phx_elph* h, int nq, const double* q3, const double* e3s, double* out) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_elph_polar_part1: null handle");
    PHX_REQUIRE(h->polar.on, "phx_elph_polar_part1: no polar parameters (phx_elph_set_polar)");
    PHX_REQUIRE(nq >= 0 && (nq == 0 || (q3 && e3s && out)), "phx_elph_polar_part1: bad argument");
    if (nq == 0) return;
    hipStream_t s = h->stream;
    const int nm = h->nm;
    DevBuf<double> dq;
    DevBuf<double2> de3, dx(size_t(nq) * nm);
    h2d_grow(dq, q3, size_t(nq) * 3, s);
    h2d_grow(de3, reinterpret_cast<const double2*>(e3s), size_t(nq) * nm * nm, s);
    launch_polar_part1(h->polar.d, nq, nm, dq.p, de3.p, dx.p, s);
    PHX_HIP(hipMemcpyAsync(out, dx.p, dx.bytes(), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
  });
}

int phx_elph_get_coupling_dims(phx_elph* h, int* d) {
  return guard([&] {
    PHX_REQUIRE(h && d, "null argument");
    d[0] = h->nRe;
    d[1] = h->nRp;
    d[2] = h->nm;
    d[3] = h->nw;
    d[4] = h->nw;
  });
}

// Device memory the reference accounts for an InteractionElPhWan
// (getDeviceMemoryUsage, interaction_elph.cpp:703-707): the coupling tensor
// and elPhCached (nRp nm nb1 nw complex once cacheElPh has run) at 16 bytes
// per element, the R vectors and their degeneracies at 8.
static double elph_accounted_bytes(const phx_elph* h) {
  const double cached = h->nb1 > 0 ? double(h->nRp) * h->nm * h->nb1 * h->nw : 0.;
  return 16. * (double(h->nRe) * h->nRp * h->nm * h->nw * h->nw + cached) +
         8. * (4. * h->nRp + 4. * h->nRe);
}

int phx_elph_device_memory_usage(phx_elph* h, double* bytes) {
  return guard([&] {
    PHX_REQUIRE(h && bytes, "phx_elph_device_memory_usage: null argument");
    *bytes = elph_accounted_bytes(h);
  });
}

int phx_elph_estimate_num_batches(phx_elph* h, int nk2, int nb1, double available) {
  int result = 0;
  int rc = guard([&] {
    PHX_REQUIRE(h && nk2 >= 0 && nb1 >= 1, "phx_elph_estimate_num_batches: bad argument");
    if (available <= 0) {
      // a DeviceManager (common_kokkos.cpp:104-141) holding only this object
      double total = 16.e9;
      if (const char* e = std::getenv("MAXMEM")) total = std::atof(e) * 1.e9;
      available = total - elph_accounted_bytes(h);
    }
    // memory model of interaction_elph.cpp:459-474 (numElBands = nw, numPhBands = nm)
    const double nw = h->nw, nm = h->nm, nRp = h->nRp;
    const double evs = 16 * (nw * nw + nm * nm);
    const double phase = 16 * nRp;
    const double g3 = 2 * 16 * nm * nb1 * nw;
    const double g4 = 2 * 16 * nm * nb1 * nw;
    const double gFinal = 2 * 16 * nm * nb1 * nw;
    const double coupling = 16 * nb1 * nw * nm;
    const double polar = 16 * nm * nb1 * nw;
    const double maxUsage = nk2 * (evs + polar + std::max({phase + g3, g3 + g4, g4 + gFinal, gFinal + coupling}));
    // :477 and :482-488 (the reference divides by nk2 also when it is 0: no error then)
    PHX_REQUIRE(nk2 == 0 || available >= maxUsage / nk2, "Insufficient memory!");
    result = int(std::ceil(maxUsage / available));
  });
  return rc ? rc : result;
}

int phx_elph_destroy(phx_elph* h) {
  return guard([&] { delete h; });
}

int phx_lw_setup(phx_lw** h, const phx_problem* prob, int precision, int k1_batch, int transform) {
  return guard([&] {
    PHX_REQUIRE(h && prob, "phx_lw_setup: null argument");
    auto* L = new phx_lw();
    try {
      lw_setup(L, prob, precision, k1_batch, transform);
    } catch (...) {
      delete L;
      throw;
    }
    *h = L;
  });
}

int phx_lw_setup_file(phx_lw** h, const phx_problem* prob, const char* elph_path, int precision,
                      int k1_batch, int transform) {
  return guard([&] {
    PHX_REQUIRE(h && prob && elph_path, "phx_lw_setup_file: null argument");
    phx::ElphFile f(elph_path);
    const phx::ElphHeader& hd = f.header();
    PHX_REQUIRE(hd.nw == prob->nw && hd.nm == prob->nm,
                "phx_lw_setup_file: the el-ph file's numElBands / numPhModes (" +
                    std::to_string(hd.nw) + ", " + std::to_string(hd.nm) +
                    ") differ from the band structures (" + std::to_string(prob->nw) + ", " +
                    std::to_string(prob->nm) + ")");
    std::vector<double> eR(3 * size_t(hd.nRe)), eD(size_t(hd.nRe)), pR(3 * size_t(hd.nRp)),
        pD(size_t(hd.nRp));
    f.bravais(eR.data(), eD.data(), pR.data(), pD.data());
    phx_problem q = *prob;
    q.nRe = hd.nRe;
    q.nRp = hd.nRp;
    q.Re = eR.data();
    q.degE = eD.data();
    q.Rp = pR.data();
    q.degP = pD.data();
    q.gW = nullptr;
    q.gwDecay = nullptr;
    auto* L = new phx_lw();
    try {
      lw_setup(L, &q, precision, k1_batch, transform, &f);
    } catch (...) {
      delete L;
      throw;
    }
    // the stored problem must not point at the arrays above
    L->prob.Re = L->prob.degE = L->prob.Rp = L->prob.degP = nullptr;
    *h = L;
  });
}


int phx_phel_setup(phx_phel** h, const phx_problem* prob, phx_elph* coupling, const int* q_mesh) {
  return guard([&] {
    PHX_REQUIRE(h && prob && coupling && q_mesh, "phx_phel_setup: null argument");
    *h = nullptr;
    auto* P = new phx_phel();
    try {
      phel_setup(P, *prob, coupling, q_mesh);
    } catch (...) {
      delete P;
      throw;
    }
    *h = P;
  });
}

int phx_phel_num_active(phx_phel* h) { return h ? h->na : -1; }

int phx_phel_set_timing(phx_phel* h, int on) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_phel_set_timing: null handle");
    h->timing = on != 0;
    h->cacheMs = h->coupMs = h->redMs = h->coupFlops = 0.;
    h->coupLaunches = h->cacheLaunches = 0;
  });
}

int phx_phel_kernel_times(phx_phel* h, double* out6) {
  return guard([&] {
    PHX_REQUIRE(h && out6, "phx_phel_kernel_times: null argument");
    out6[0] = h->cacheMs;
    out6[1] = h->coupMs;
    out6[2] = h->redMs;
    out6[3] = double(h->coupLaunches);
    out6[4] = double(h->cacheLaunches);
    out6[5] = h->coupFlops;
  });
}

long long phx_phel_run(phx_phel* h, int nk1, const int* k1_list) {
  long long before = 0;
  int rc = guard([&] {
    PHX_REQUIRE(h, "phx_phel_run: null handle");
    before = h->pairs;
    phel_run(h, nk1, k1_list);
  });
  return rc < 0 ? rc : h->pairs - before;
}

int phx_phel_get_linewidths(phx_phel* h, double* out, int average) {
  return guard([&] {
    PHX_REQUIRE(h && out, "phx_phel_get_linewidths: null argument");
    phel_get(h, out, average);
  });
}

int phx_phel_get_phonons(phx_phel* h, double* omega) {
  return guard([&] {
    PHX_REQUIRE(h && omega, "phx_phel_get_phonons: null argument");
    std::copy(h->omega.begin(), h->omega.end(), omega);
  });
}

int phx_phel_reset(phx_phel* h) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_phel_reset: null handle");
    launch_fill<double>(h->lw.p, h->lw.n, 0., h->s);
    PHX_HIP(hipStreamSynchronize(h->s));
    h->pairs = 0;
  });
}

int phx_phel_destroy(phx_phel* h) {
  return guard([&] { delete h; });
}

int phx_lw_num_active(phx_lw* h) { return h ? h->na : -1; }
int phx_lw_num_states(phx_lw* h) { return h ? h->numStates : -1; }

int phx_lw_get_window(phx_lw* h, int* active, int* lo, int* hi, double* energies) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    if (active) std::copy(h->active.begin(), h->active.end(), active);
    if (lo) std::copy(h->lo.begin(), h->lo.end(), lo);
    if (hi) std::copy(h->hi.begin(), h->hi.end(), hi);
    if (energies) std::copy(h->stateE.begin(), h->stateE.end(), energies);
  });
}

int phx_lw_get_phonons(phx_lw* h, double* omega, unsigned char* inSet) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    if (omega) std::copy(h->omegaAll.begin(), h->omegaAll.end(), omega);
    if (inSet) std::copy(h->qInSet.begin(), h->qInSet.end(), inSet);
  });
}

long long phx_lw_run(phx_lw* h, int nk1, const int* k1list, int k2b, int k2e) {
  long long pairs = 0;
  int rc = guard([&] {
    PHX_REQUIRE(h, "null handle");
    pairs = lw_run(h, nk1, k1list, k2b, k2e);
  });
  return rc ? rc : pairs;
}

long long phx_lw_plan(phx_lw* h, int nk1, const int* k1_list, int k2_begin, int k2_end) {
  int rc = guard([&] {
    PHX_REQUIRE(h, "null handle");
    lw_plan(h, nk1, k1_list, k2_begin, k2_end);
  });
  return rc < 0 ? rc : (long long)h->plan.size();
}

long long phx_lw_plan_run(phx_lw* h, int graph) {
  long long pairs = 0;
  int rc = guard([&] {
    PHX_REQUIRE(h, "null handle");
    pairs = lw_plan_run(h, graph);
  });
  return rc < 0 ? rc : pairs;
}

// Autotuner of the batch-level tile shapes: k1 batch (GEMM-form M tile), mesh
// lines per line batch, panels per separable launch.  Coordinate descent on
// the wall time of full runs over the sample k1_list; the handle's linewidths
// are restored afterwards.  out4 = {k1_batch, lines_per_batch, panel_chunk, best seconds}.
static void lw_autotune(LwHandle* L, int nk1, const int* k1list, double* out4) {
  hipStream_t s = L->stream;
  PHX_HIP(hipStreamSynchronize(s));
  DevBuf<double> keep(L->lw.n);
  PHX_HIP(hipMemcpyAsync(keep.p, L->lw.p, L->lw.bytes(), hipMemcpyDeviceToDevice, s));
  auto timeit = [&]() {
    PHX_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    lw_run(L, nk1, k1list, 0, -1);
    PHX_HIP(hipStreamSynchronize(s));
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  timeit();  // buffers grown to size
  double best = timeit();
  auto tune = [&](int* knob, std::initializer_list<int> values) {
    int bestv = *knob;
    for (int v : values) {
      if (v == bestv) continue;
      *knob = v;
      const double t = timeit();
      if (t < best) {
        best = t;
        bestv = v;
      }
    }
    *knob = bestv;
  };
  tune(&L->k1Batch, {8, 16, 32});
  if (L->lineMin > 0) tune(&L->linesPerBatch, {1, 2, 4, 8, 16});
  if (L->transform == PHX_TRANSFORM_SEPARABLE) tune(&L->panelChunk, {16, 48, 128});
  PHX_HIP(hipMemcpyAsync(L->lw.p, keep.p, L->lw.bytes(), hipMemcpyDeviceToDevice, s));
  PHX_HIP(hipStreamSynchronize(s));
  out4[0] = L->k1Batch;
  out4[1] = L->linesPerBatch;
  out4[2] = L->panelChunk;
  out4[3] = best;
}

int phx_lw_autotune(phx_lw* h, int nk1, const int* k1_list, double* out4) {
  return guard([&] {
    PHX_REQUIRE(h && out4, "null argument");
    PHX_REQUIRE(!h->lsCollect, "phx_lw_autotune: not while collecting screen samples");
    lw_autotune(h, nk1, k1_list, out4);
  });
}

int phx_lw_set_batching(phx_lw* h, int k1_batch, int lines_per_batch) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_REQUIRE(k1_batch >= 1 && k1_batch <= kCacheNKK && lines_per_batch >= 1 &&
                    lines_per_batch <= kLinesPerBatch,
                "phx_lw_set_batching: k1_batch in 1..32, lines_per_batch in 1..16");
    h->k1Batch = k1_batch;
    h->linesPerBatch = lines_per_batch;
  });
}

int phx_lw_get_linewidths(phx_lw* h, double* out, int average) {
  return guard([&] {
    PHX_REQUIRE(h && out, "null argument");
    lw_get(h, out, average);
  });
}

int phx_lw_reset(phx_lw* h) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    launch_fill<double>(h->lw.p, h->lw.n, 0., h->stream);
    PHX_HIP(hipStreamSynchronize(h->stream));
  });
}

int phx_lw_set_polar(phx_lw* h, const phx_polar* polar) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_lw_set_polar: null handle");
    if (!polar) {
      h->polar.set(nullptr, h->stream);
      h->polX.release();
      return;
    }
    PHX_REQUIRE(h->transform == PHX_TRANSFORM_DIRECT,
                "phx_lw_set_polar: the polar term runs with the direct transform");
    PHX_REQUIRE(3 * polar->natoms == h->nm, "phx_lw_set_polar: nm must be 3 natoms");
    hipStream_t s = h->stream;
    h->polar.set(polar, s);
    // polarCorrectionPart1 at every mesh q (folded wavevector and its mass-scaled
    // eigenvectors, helper_el_scattering.cpp:170-195); points outside the
    // q-set have zero eigenvectors and get x = 0
    DevBuf<double> dq;
    h2d(dq, h->kcart, s);
    h->polX.alloc(size_t(h->nk) * h->nm);
    launch_polar_part1(h->polar.d, h->nk, h->nm, dq.p, h->E3all.p, h->polX.p, s);
    // q = 0 only for k2 = k1, where the reference skips the term (:289)
    launch_fill<double>(reinterpret_cast<double*>(h->polX.p), 2 * size_t(h->nm), 0., s);
    PHX_HIP(hipStreamSynchronize(s));
  });
}

int phx_lw_set_screen(phx_lw* h, double nsigma) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_REQUIRE(nsigma == 0. || nsigma >= 6., "phx_lw_set_screen: use 0 (off) or >= 6 sigma");
    PHX_REQUIRE(nsigma == 0. || !h->adaptive,
                "phx_lw_set_screen: the energy screen needs a fixed Gaussian width (adaptive smearing is on)");
    PHX_REQUIRE(nsigma == 0. || h->transform == PHX_TRANSFORM_SEPARABLE,
                "phx_lw_set_screen: the screen runs with the separable transform");
    h->screen = nsigma;
  });
}

int phx_lw_set_line_batching(phx_lw* h, int min_points) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_REQUIRE(min_points >= 0, "phx_lw_set_line_batching: min_points must be >= 0");
    PHX_REQUIRE(min_points == 0 || h->reLattice,
                "phx_lw_set_line_batching: the electronic R vectors are not lattice vectors");
    h->lineMin = min_points;
  });
}

int phx_lw_set_split_last_stage(phx_lw* h, int on) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_lw_set_split_last_stage: null handle");
    PHX_REQUIRE(on >= 0 && on <= 3, "phx_lw_set_split_last_stage: 0 to 3");
    h->splitLast = on;
  });
}

int phx_lw_set_epilogue_overlap(phx_lw* h, int on) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_lw_set_epilogue_overlap: null handle");
    PHX_REQUIRE(on == 0 || on == 1, "phx_lw_set_epilogue_overlap: 0 or 1");
    h->epOverlap = on;
  });
}

int phx_lw_set_fused_y_stage(phx_lw* h, int on) {
  return guard([&] {
    PHX_REQUIRE(h, "phx_lw_set_fused_y_stage: null handle");
    PHX_REQUIRE(on == 0 || on == 1, "phx_lw_set_fused_y_stage: 0 or 1");
    h->fuseYX = on;
  });
}

int phx_lw_last_stage_fused_y(phx_lw* h, int* on) {
  return guard([&] {
    PHX_REQUIRE(h && on, "null argument");
    PHX_HIP(hipStreamSynchronize(h->stream));
    lw_harvest_timing(h);
    *on = h->yxLaunches > 0 && h->yxLaunches == h->splitLaunches ? 1 : 0;
  });
}

int phx_lw_set_panel_chunk(phx_lw* h, int panels) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_REQUIRE(panels >= 0, "phx_lw_set_panel_chunk: panels must be >= 0 (0: automatic)");
    h->panelChunk = panels;
  });
}

// ---- learned screen: training on the collected samples
static void ls_alloc(LwHandle* L) {
  hipStream_t s = L->stream;
  if (!L->mlpP.p) {
    L->mlpP.alloc(kMlpParams);
    L->mlpG.alloc(kMlpParams);
    L->mlpM1.alloc(kMlpParams);
    L->mlpM2.alloc(kMlpParams);
    L->mlpLoss.alloc(4);
  }
  if (!L->lsCounts.p) {
    L->lsCounts.alloc(4);
    L->lsStats.alloc(2);
    L->lsFail.alloc(1);
    launch_fill<int>(reinterpret_cast<int*>(L->lsCounts.p), 8, 0, s);
    launch_fill<double>(L->lsStats.p, 2, 0., s);
    launch_fill<int>(L->lsFail.p, 1, 0, s);
  }
}

static void ls_train(LwHandle* L, int epochs, int batch, float lr, uint64_t seed) {
  hipStream_t s = L->stream;
  const int n = int(L->trT.size());
  PHX_REQUIRE(n >= 16, "learned screen: collect at least 16 training samples first");
  PHX_REQUIRE(epochs >= 1 && batch >= 1 && lr > 0.f, "learned screen: bad training settings");
  ls_alloc(L);
  double mean = 0., var = 0.;
  for (float t : L->trT) mean += t;
  mean /= n;
  for (float t : L->trT) var += (t - mean) * (t - mean);
  const double sd = std::max(std::sqrt(var / n), 1e-6);
  std::vector<float> tn(n);
  for (int i = 0; i < n; ++i) tn[i] = float((L->trT[i] - mean) / sd);
  // Glorot-uniform weights, zero biases
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-1., 1.);
  std::vector<float> P(kMlpParams, 0.f);
  auto init = [&](int off, int fin, int fout) {
    const double a = std::sqrt(6. / (fin + fout));
    for (int i = 0; i < fin * fout; ++i) P[off + i] = float(a * U(rng));
  };
  init(kMlpOffW1, kMlpIn, kMlpHid);
  init(kMlpOffW2, kMlpHid, kMlpHid);
  init(kMlpOffW3, kMlpHid, 1);
  PHX_HIP(hipMemcpyAsync(L->mlpP.p, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice, s));
  launch_fill<float>(L->mlpM1.p, kMlpParams, 0.f, s);
  launch_fill<float>(L->mlpM2.p, kMlpParams, 0.f, s);
  L->trXd.reserve(size_t(n) * kMlpIn);
  L->trTd.reserve(n);
  L->mlpScratch.reserve(mlp_scratch_floats(std::max(n, batch)));
  std::vector<float> Xp(size_t(n) * kMlpIn), Tp(n);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  auto full_loss = [&]() {
    PHX_HIP(hipMemcpyAsync(L->trXd.p, L->trX.data(), size_t(n) * kMlpIn * sizeof(float),
                           hipMemcpyHostToDevice, s));
    PHX_HIP(hipMemcpyAsync(L->trTd.p, tn.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    mlp_backward(L->mlpP.p, L->trXd.p, L->trTd.p, n, L->mlpScratch.p, L->mlpG.p, L->mlpLoss.p, s);
    float v = 0.f;
    PHX_HIP(hipMemcpyAsync(&v, L->mlpLoss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    PHX_HIP(hipStreamSynchronize(s));
    return double(v);
  };
  L->lsLoss[0] = full_loss();
  int step = 0;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(perm.begin(), perm.end(), rng);
    for (int i = 0; i < n; ++i) {
      std::copy(L->trX.begin() + size_t(perm[i]) * kMlpIn, L->trX.begin() + size_t(perm[i] + 1) * kMlpIn,
                Xp.begin() + size_t(i) * kMlpIn);
      Tp[i] = tn[perm[i]];
    }
    PHX_HIP(hipMemcpyAsync(L->trXd.p, Xp.data(), Xp.size() * sizeof(float), hipMemcpyHostToDevice, s));
    PHX_HIP(hipMemcpyAsync(L->trTd.p, Tp.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    for (int b = 0; b < n; b += batch) {
      const int nb = std::min(batch, n - b);
      mlp_backward(L->mlpP.p, L->trXd.p + size_t(b) * kMlpIn, L->trTd.p + b, nb, L->mlpScratch.p,
                   L->mlpG.p, L->mlpLoss.p, s);
      mlp_adam(L->mlpP.p, L->mlpG.p, L->mlpM1.p, L->mlpM2.p, ++step, lr, s);
    }
    PHX_HIP(hipStreamSynchronize(s));  // the host staging arrays are refilled next epoch
  }
  L->lsLoss[1] = full_loss();
  {
    // under-prediction margin: ln-residuals target - prediction (standardized units)
    std::vector<float> yv(n);
    PHX_HIP(hipMemcpy(yv.data(), L->mlpScratch.p + size_t(n) * (4 * kMlpHid + 1), n * sizeof(float),
                      hipMemcpyDeviceToHost));
    // over the samples the screen may skip (no row within 3 sigma: feature 11
    // is zero), all samples when fewer than 100 of them
    std::vector<double> r;
    for (int pass = 0; pass < 2 && r.size() < 100; ++pass) {
      r.clear();
      for (int i = 0; i < n; ++i)
        if (pass == 1 || L->trX[size_t(i) * kMlpIn + 11] == 0.f) r.push_back((double(tn[i]) - yv[i]) * sd);
    }
    std::sort(r.begin(), r.end());
    const int nr = int(r.size());
    L->lsMargin = std::max(0., r[std::min(nr - 1, int(0.999 * nr))]);
  }
  // fold the target normalization into the output layer: the network predicts ln c
  PHX_HIP(hipMemcpy(P.data(), L->mlpP.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < kMlpHid; ++i) P[kMlpOffW3 + i] = float(P[kMlpOffW3 + i] * sd);
  P[kMlpOffB3] = float(P[kMlpOffB3] * sd + mean);
  PHX_HIP(hipMemcpy(L->mlpP.p, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  L->mlpTrained = true;
}

int phx_lw_screen_collect(phx_lw* h, int nk1, const int* k1_list, double sample_frac,
                          unsigned long long seed) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_REQUIRE(h->transform == PHX_TRANSFORM_SEPARABLE && h->screen > 0.,
                "learned screen: needs the separable transform with the energy screen on");
    PHX_REQUIRE(sample_frac > 0. && sample_frac <= 1., "learned screen: sample fraction in (0, 1]");
    ls_alloc(h);
    h->lsSample = sample_frac;
    h->lsSeed = seed;
    h->lsCollect = true;
    try {
      lw_run(h, nk1, k1_list, 0, -1);
      PHX_HIP(hipStreamSynchronize(h->stream));
    } catch (...) {
      h->lsCollect = false;
      throw;
    }
    h->lsCollect = false;
  }) < 0 ? -1 : int(h->trT.size());
}

int phx_lw_screen_train(phx_lw* h, int epochs, int batch, double lr, unsigned long long seed,
                        double* loss_first_last) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    ls_train(h, epochs, batch, float(lr), seed);
    if (loss_first_last) {
      loss_first_last[0] = h->lsLoss[0];
      loss_first_last[1] = h->lsLoss[1];
    }
  });
}

int phx_lw_set_learned_screen(phx_lw* h, int on, double threshold, double audit_frac,
                              unsigned long long seed) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_REQUIRE(!on || h->mlpTrained, "learned screen: train the network first");
    PHX_REQUIRE(threshold > 0. && threshold < 1. && audit_frac >= 0. && audit_frac <= 1.,
                "learned screen: threshold in (0, 1), audit fraction in [0, 1]");
    ls_alloc(h);
    h->lsMode = on ? 1 : 0;
    h->lsThr = threshold;
    h->lsAudit = audit_frac;
    h->lsSeed = seed;
  });
}

int phx_lw_screen_stats(phx_lw* h, double* out8) {
  return guard([&] {
    PHX_REQUIRE(h && out8, "null argument");
    for (int i = 0; i < 8; ++i) out8[i] = 0.;
    if (h->lsCounts.p) {
      long long c[4];
      double st[2];
      PHX_HIP(hipStreamSynchronize(h->stream));
      PHX_HIP(hipMemcpy(c, h->lsCounts.p, sizeof(c), hipMemcpyDeviceToHost));
      PHX_HIP(hipMemcpy(st, h->lsStats.p, sizeof(st), hipMemcpyDeviceToHost));
      for (int i = 0; i < 4; ++i) out8[i] = double(c[i]);
      out8[4] = st[0];
      out8[5] = c[2] > 0 ? st[1] / double(c[2]) : 0.;
    }
    out8[6] = double(h->trT.size());
    out8[7] = h->lsMargin;
  });
}

int phx_lw_screen_reset_stats(phx_lw* h) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    ls_alloc(h);
    launch_fill<int>(reinterpret_cast<int*>(h->lsCounts.p), 8, 0, h->stream);
    launch_fill<double>(h->lsStats.p, 2, 0., h->stream);
    PHX_HIP(hipStreamSynchronize(h->stream));
  });
}

int phx_lw_screen_params(phx_lw* h, float* params, int set) {
  return guard([&] {
    PHX_REQUIRE(h && params, "null argument");
    ls_alloc(h);
    if (set) {
      PHX_HIP(hipMemcpy(h->mlpP.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
      h->mlpTrained = true;
    } else {
      PHX_REQUIRE(h->mlpTrained, "learned screen: no trained network");
      PHX_HIP(hipMemcpy(params, h->mlpP.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    }
  });
}

int phx_mlp_num_params(void) { return kMlpParams; }

int phx_mlp_gradients(const float* params, const float* X, const float* t, int n, float* grads,
                      float* loss, float* y) {
  return guard([&] {
    PHX_REQUIRE(params && X && t && n > 0, "phx_mlp_gradients: bad arguments");
    hipStream_t s = 0;
    DevBuf<float> P, Xd, Td, G, sc, l;
    h2d(P, params, kMlpParams, s);
    h2d(Xd, X, size_t(n) * kMlpIn, s);
    h2d(Td, t, size_t(n), s);
    G.alloc(kMlpParams);
    sc.alloc(mlp_scratch_floats(n));
    l.alloc(1);
    mlp_backward(P.p, Xd.p, Td.p, n, sc.p, G.p, l.p, s);
    PHX_HIP(hipStreamSynchronize(s));
    if (grads) PHX_HIP(hipMemcpy(grads, G.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) PHX_HIP(hipMemcpy(loss, l.p, sizeof(float), hipMemcpyDeviceToHost));
    // y (network output) sits after H1, H2, dZ2, dZ1 and dy in the scratch
    if (y) PHX_HIP(hipMemcpy(y, sc.p + size_t(n) * (4 * kMlpHid + 1), n * sizeof(float),
                             hipMemcpyDeviceToHost));
  });
}

int phx_lw_set_timing(phx_lw* h, int on) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_HIP(hipStreamSynchronize(h->stream));
    lw_harvest_timing(h);
    if (on && !h->ev[0])
      for (auto& e : h->ev) PHX_HIP(hipEventCreate(&e));
    h->timing = on != 0;
    h->kernelMs[0] = h->kernelMs[1] = h->kernelMs[2] = 0.;
    h->kernelLaunches = 0;
    h->couplingFlops = 0.;
    h->xMs = h->xgMs = h->epMs = 0.;
    h->xLaunches = h->splitLaunches = 0;
    h->yxT1Bytes = 0.;
    h->yxLaunches = 0;
    h->xevUsed = 0;
    h->xevSplit.clear();
    if (!h->xKept.p) h->xKept.alloc(1);
    if (!h->xBands.p) h->xBands.alloc(1);
    if (!h->xSlots.p) h->xSlots.alloc(1);
    launch_fill<int>(reinterpret_cast<int*>(h->xKept.p), 2, 0, h->stream);
    launch_fill<int>(reinterpret_cast<int*>(h->xBands.p), 2, 0, h->stream);
    launch_fill<int>(reinterpret_cast<int*>(h->xSlots.p), 2, 0, h->stream);
    PHX_HIP(hipStreamSynchronize(h->stream));
  });
}

int phx_lw_kernel_times(phx_lw* h, double* out5) {
  return guard([&] {
    PHX_REQUIRE(h && out5, "null argument");
    PHX_HIP(hipStreamSynchronize(h->stream));
    lw_harvest_timing(h);
    out5[0] = h->kernelMs[0];
    out5[1] = h->kernelMs[1];
    out5[2] = h->kernelMs[2];
    out5[3] = double(h->kernelLaunches);
    out5[4] = h->couplingFlops;
  });
}

int phx_lw_stage_times(phx_lw* h, double* out4) {
  return guard([&] {
    PHX_REQUIRE(h && out4, "null argument");
    PHX_HIP(hipStreamSynchronize(h->stream));
    lw_harvest_timing(h);
    unsigned long long kept = 0;
    if (h->xKept.p) PHX_HIP(hipMemcpy(&kept, h->xKept.p, sizeof(kept), hipMemcpyDeviceToHost));
    out4[0] = h->xMs;
    out4[1] = double(h->xLaunches);
    out4[2] = double(kept);
    // complex multiply-adds of the last-stage transform for the rows computed
    out4[3] = 8. * double(kept) * h->nG0 * double(h->nm) * h->nw;
  });
}

int phx_lw_last_stage_split_times(phx_lw* h, double* out4) {
  return guard([&] {
    PHX_REQUIRE(h && out4, "null argument");
    PHX_HIP(hipStreamSynchronize(h->stream));
    lw_harvest_timing(h);
    unsigned long long bands = 0;
    if (h->xBands.p) PHX_HIP(hipMemcpy(&bands, h->xBands.p, sizeof(bands), hipMemcpyDeviceToHost));
    out4[0] = h->xgMs;
    out4[1] = h->epMs;
    out4[2] = double(h->splitLaunches);
    out4[3] = double(bands);
  });
}

int phx_lw_last_stage_split_bytes(phx_lw* h, double* out2) {
  return guard([&] {
    PHX_REQUIRE(h && out2, "null argument");
    PHX_HIP(hipStreamSynchronize(h->stream));
    lw_harvest_timing(h);
    unsigned long long slots = 0, rows = 0, bands = 0;
    if (h->xSlots.p) PHX_HIP(hipMemcpy(&slots, h->xSlots.p, sizeof(slots), hipMemcpyDeviceToHost));
    if (h->xKept.p) PHX_HIP(hipMemcpy(&rows, h->xKept.p, sizeof(rows), hipMemcpyDeviceToHost));
    if (h->xBands.p) PHX_HIP(hipMemcpy(&bands, h->xBands.p, sizeof(bands), hipMemcpyDeviceToHost));
    // complex FP32: the x-stage reads nG0 x (nm nw) per slot and writes nm nw per kept row;
    // with the y-stage inside (k_sep_yx) it reads T1 instead of the slabs (slots not counted)
    const double row = 8. * h->nm * h->nw;
    out2[0] = double(slots) * h->nG0 * row + h->yxT1Bytes;
    out2[1] = h->splitLaunches > 0 ? double(rows) * row : 0.;
    if (h->splitLast == 3 && h->splitLaunches > 0) {
      // rotation in the x-stage: it reads U2 (nw complex FP64 per kept band) and writes
      // g' (nm complex FP64 per kept band), which the epilogue reads back
      const double gp = 16. * double(bands) * h->nm;
      out2[0] += gp + 16. * double(bands) * h->nw;
      out2[1] = gp;
    } else {
      out2[0] += out2[1];
    }
  });
}

__global__ void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

int phx_stream_copy(unsigned long long bytes, int reps, double* ms_per_copy) {
  return guard([&] {
    PHX_REQUIRE(bytes >= 16 && reps >= 1 && ms_per_copy, "phx_stream_copy: bad arguments");
    const size_t n = size_t(bytes / 16);
    DevBuf<float4> a(n), b(n);
    hipStream_t s = 0;
    launch_fill<float>(reinterpret_cast<float*>(a.p), 4 * n, 1.f, s);
    hipEvent_t e0, e1;
    PHX_HIP(hipEventCreate(&e0));
    PHX_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(256), 0, s, a.p, b.p, n);
    PHX_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(256), 0, s, a.p, b.p, n);
    PHX_HIP(hipEventRecord(e1, s));
    PHX_HIP(hipEventSynchronize(e1));
    PHX_HIP(hipGetLastError());
    float ms = 0.f;
    PHX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_copy = double(ms) / reps;
  });
}

int phx_lw_sync(phx_lw* h) {
  return guard([&] {
    PHX_REQUIRE(h, "null handle");
    PHX_HIP(hipStreamSynchronize(h->stream));
  });
}

void* phx_lw_stream(phx_lw* h) { return h ? static_cast<void*>(h->stream) : nullptr; }

int phx_lw_copy_gw(phx_lw* h, double* out) {
  return guard([&] {
    PHX_REQUIRE(h && out, "null argument");
    PHX_REQUIRE(h->precision == PHX_FP64, "phx_lw_copy_gw: only FP64 handles keep the FP64 tensor");
    PHX_HIP(hipStreamSynchronize(h->stream));
    PHX_HIP(hipMemcpy(out, h->gW.p, h->gW.bytes(), hipMemcpyDeviceToHost));
  });
}

int phx_lw_destroy(phx_lw* h) {
  return guard([&] { delete h; });
}

int phx_synth_gw_host(double* gW, int nw, int nm, int nRp, int nRe, const double* decay,
                      unsigned long long seed) {
  return guard([&] {
    PHX_REQUIRE(gW && decay, "null argument");
    synth_gw_host(gW, size_t(nRe) * nRp * nm * nw * nw, size_t(nm) * nw * nw, decay, seed);
  });
}

}  // extern "C"
