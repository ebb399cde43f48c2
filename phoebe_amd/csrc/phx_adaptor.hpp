// Header-only C++ adaptor with the call signatures of Phoebe's
// InteractionElPhWan (src/interaction/interaction_elph.h:30-219) and
// kokkosZHEEV (src/common_kokkos.h:49) on top of the C ABI in
// include/phx_elph.h.
//
// Templated on the matrix / tensor types so that it compiles against Eigen
// inside Phoebe (Eigen::MatrixXcd, Eigen::Vector3d, Eigen::Tensor<double,3>,
// Eigen::VectorXi, Eigen::VectorXcd; the constructor takes
// Eigen::Tensor<std::complex<double>,5>, Eigen::MatrixXd, Eigen::VectorXd,
// and Phoebe's Crystal, PhononH0 and Context), or against any types with the
// same column-major storage and accessors: data(), rows()/cols() or size(),
// dimension(i) for tensors, operator()(i, j[, k]) for the Crystal / PhononH0
// fields of the polar term.  Errors follow the reference's Error() behaviour
// (exceptions.cpp:6-16): print and terminate, unless the caller installs its
// own handler.
#pragma once

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/phx_elph.h"

namespace phx_adaptor {

using ErrorHandler = std::function<void(const std::string&)>;

inline ErrorHandler& error_handler() {
  static ErrorHandler h = [](const std::string& m) {
    std::fprintf(stderr, "Error: %s\n", m.c_str());
    std::exit(1);
  };
  return h;
}

inline void check(int rc) {
  if (rc < 0) error_handler()(phx_last_error());
}

// InteractionElPhWan on the GPU.  Matrix: column-major std::complex<double>
// matrix (rows(), cols(), size(), data()); Vec3: 3-vector with data();
// Tensor3: rank-3 double tensor with the constructor Tensor3(nb1, nb2, nm)
// and column-major data(); VectorXi: int vector constructible with a size;
// VectorXcd: complex vector constructible with a size.
template <class Matrix, class Vec3, class Tensor3, class VectorXi = std::vector<int>,
          class VectorXcd = std::vector<std::complex<double>>>
class InteractionElPhWan {
 public:
  // interaction_elph.h:93-100 / interaction_elph.cpp:10-62.  couplingWannier
  // is the reference's tensor (nw, nw, nm, nRp, nRe) column-major;
  // elBravaisVectors / phBravaisVectors are (3, nR) column-major Cartesian.
  // The polar term is used when phononH0 carries a dielectric tensor and the
  // crystal has more than one species (:24-32).
  template <class Crystal, class Tensor5, class MatrixXd, class VectorXd, class PhononH0 = void>
  InteractionElPhWan(Crystal& crystal, const Tensor5& couplingWannier, const MatrixXd& elBravaisVectors,
                     const VectorXd& elBravaisVectorsDegeneracies, const MatrixXd& phBravaisVectors,
                     const VectorXd& phBravaisVectorsDegeneracies, PhononH0* phononH0 = nullptr,
                     int precision = PHX_FP64)
      : nw_(int(couplingWannier.dimension(0))), nm_(int(couplingWannier.dimension(2))) {
    check(phx_elph_create(&h_, reinterpret_cast<const double*>(couplingWannier.data()), nw_, nm_,
                          int(couplingWannier.dimension(3)), int(couplingWannier.dimension(4)),
                          elBravaisVectors.data(), elBravaisVectorsDegeneracies.data(), phBravaisVectors.data(),
                          phBravaisVectorsDegeneracies.data(), precision));
    setup_polar(crystal, phononH0);
  }

  // Flat-array form: gW as a flat complex array in the same layout.
  InteractionElPhWan(const std::complex<double>* gW, int nw, int nm, int nRp, int nRe, const double* elR,
                     const double* elDeg, const double* phR, const double* phDeg, int precision = PHX_FP64)
      : nw_(nw), nm_(nm) {
    check(phx_elph_create(&h_, reinterpret_cast<const double*>(gW), nw, nm, nRp, nRe, elR, elDeg, phR, phDeg,
                          precision));
  }

  // interaction_elph.h:182-183 / interaction_elph_parsing.cpp:672-692: the
  // interaction read from context.getElphFileName(), streamed into HBM.
  template <class Context, class Crystal, class PhononH0 = void>
  static InteractionElPhWan parse(Context& context, Crystal& crystal, PhononH0* phononH0 = nullptr,
                                  int precision = PHX_FP64) {
    InteractionElPhWan out;
    const std::string path = context.getElphFileName();
    check(phx_elph_create_from_file(&out.h_, path.c_str(), precision));
    int d[5];
    check(phx_elph_get_coupling_dims(out.h_, d));
    out.nm_ = d[2];
    out.nw_ = d[3];
    out.setup_polar(crystal, phononH0);
    return out;
  }

  InteractionElPhWan(const InteractionElPhWan&) = delete;
  InteractionElPhWan& operator=(const InteractionElPhWan&) = delete;
  InteractionElPhWan(InteractionElPhWan&& o) noexcept { swap(o); }
  InteractionElPhWan& operator=(InteractionElPhWan&& o) noexcept {
    if (this != &o) swap(o);
    return *this;
  }
  ~InteractionElPhWan() {
    if (h_) phx_elph_destroy(h_);
  }

  // interaction_elph.h:161
  void cacheElPh(const Matrix& eigvec1, const Vec3& k1C) {
    check(phx_elph_cache(h_, reinterpret_cast<const double*>(eigvec1.data()), int(eigvec1.cols()), k1C.data()));
  }

  // interaction_elph.h:138-141 / interaction_elph.cpp:138-151
  VectorXcd polarCorrectionPart1(const Vec3& q3, const Matrix& ev3) {
    VectorXcd x(nm_);
    check(phx_elph_polar_part1(h_, 1, q3.data(), reinterpret_cast<const double*>(ev3.data()),
                               reinterpret_cast<double*>(x.data())));
    return x;
  }

  // interaction_elph.h:149-154: polarData is used only when the polar term
  // is on, as in the reference (:285-312)
  template <class VecC>
  void calcCouplingSquared(const Matrix& eigvec1, const std::vector<Matrix>& eigvecs2,
                           const std::vector<Matrix>& eigvecs3, const std::vector<Vec3>& q3Cs,
                           const std::vector<VecC>& polarData) {
    const int nq = int(eigvecs2.size());
    const int nb1 = int(eigvec1.cols());
    std::vector<int> nb2s(nq);
    std::vector<std::complex<double>> U2, e3, pol;
    std::vector<double> q3(3 * size_t(nq));
    size_t nout = 0;
    for (int i = 0; i < nq; ++i) {
      nb2s[i] = int(eigvecs2[i].cols());
      U2.insert(U2.end(), eigvecs2[i].data(), eigvecs2[i].data() + eigvecs2[i].size());
      e3.insert(e3.end(), eigvecs3[i].data(), eigvecs3[i].data() + eigvecs3[i].size());
      for (int d = 0; d < 3; ++d) q3[3 * i + d] = q3Cs[i].data()[d];
      nout += size_t(nb1) * nb2s[i] * nm_;
    }
    if (usePolarCorrection_) {
      if (int(polarData.size()) != nq) error_handler()("calcCouplingSquared: polarData needed for every q3");
      for (const auto& p : polarData) pol.insert(pol.end(), p.data(), p.data() + nm_);
    }
    std::vector<double> out(nout);
    check(phx_elph_coupling_sq(h_, nq, reinterpret_cast<const double*>(U2.data()), nb2s.data(),
                               reinterpret_cast<const double*>(e3.data()), q3.data(),
                               usePolarCorrection_ ? reinterpret_cast<const double*>(pol.data()) : nullptr,
                               out.data()));
    cache_.clear();
    size_t off = 0;
    for (int i = 0; i < nq; ++i) {
      Tensor3 t(nb1, nb2s[i], nm_);
      const size_t n = size_t(nb1) * nb2s[i] * nm_;
      std::copy(out.begin() + off, out.begin() + off + n, t.data());
      cache_.push_back(std::move(t));
      off += n;
    }
  }

  // interaction_elph.h:174
  Tensor3& getCouplingSquared(const int& ik2) { return cache_[ik2]; }

  // interaction_elph.h:197: the device extents (nRe, nRp, nm, nw, nw) as the
  // reference returns them (interaction_elph.cpp:450-456)
  VectorXi getCouplingDimensions() {
    int d[5];
    check(phx_elph_get_coupling_dims(h_, d));
    VectorXi v(5);
    for (int i = 0; i < 5; ++i) v.data()[i] = d[i];
    return v;
  }

  // interaction_elph.h:205.  availableMemory: the caller's
  // kokkosDeviceMemory->getAvailableMemory() (set with setAvailableMemory);
  // unset, a device manager of MAXMEM holding only this object
  int estimateNumBatches(const int& nk2, const int& nb1) {
    int n = phx_elph_estimate_num_batches(h_, nk2, nb1, availableMemory_);
    check(n);
    return n;
  }
  void setAvailableMemory(double bytes) { availableMemory_ = bytes; }

  // interaction_elph.cpp:703-707
  double getDeviceMemoryUsage() {
    double b = 0.;
    check(phx_elph_device_memory_usage(h_, &b));
    return b;
  }

  bool usePolarCorrection() const { return usePolarCorrection_; }

 private:
  InteractionElPhWan() = default;

  void swap(InteractionElPhWan& o) noexcept {
    std::swap(h_, o.h_);
    std::swap(nw_, o.nw_);
    std::swap(nm_, o.nm_);
    std::swap(usePolarCorrection_, o.usePolarCorrection_);
    std::swap(availableMemory_, o.availableMemory_);
    std::swap(cache_, o.cache_);
  }

  template <class Crystal, class PhononH0>
  void setup_polar(Crystal& crystal, PhononH0* phononH0) {
    if constexpr (!std::is_void<PhononH0>::value) {
      if (phononH0 == nullptr) return;
      const auto epsilon = phononH0->getDielectricMatrix();
      double eps[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) eps[3 * i + j] = epsilon(i, j);
      if (!phx_polar_in_use(eps, int(crystal.getNumSpecies()))) return;
      // the fields polarCorrectionPart1 gathers (interaction_elph.cpp:140-151)
      const auto positions = crystal.getAtomicPositions();  // (natoms, 3), Bohr
      const auto born = phononH0->getBornCharges();         // (natoms, 3, 3)
      const auto recip = crystal.getReciprocalUnitCell();   // columns b_i
      const auto coarse = phononH0->getCoarseGrid();
      const int natoms = int(positions.rows());
      std::vector<double> pos(3 * size_t(natoms)), z(9 * size_t(natoms));
      for (int a = 0; a < natoms; ++a)
        for (int i = 0; i < 3; ++i) {
          pos[3 * a + i] = positions(a, i);
          for (int j = 0; j < 3; ++j) z[9 * a + 3 * i + j] = born(a, i, j);
        }
      phx_polar pp{};
      pp.natoms = natoms;
      pp.atomPos = pos.data();
      pp.bornCharges = z.data();
      pp.dielectric = eps;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) pp.recip[i + 3 * j] = recip(i, j);
      pp.volume = crystal.getVolumeUnitCell();
      for (int i = 0; i < 3; ++i) pp.coarseMesh[i] = int(coarse(i));
      check(phx_elph_set_polar(h_, &pp));
      usePolarCorrection_ = true;
    } else {
      (void)crystal;
      (void)phononH0;
    }
  }

  phx_elph* h_ = nullptr;
  int nw_ = 0, nm_ = 0;
  bool usePolarCorrection_ = false;
  double availableMemory_ = 0.;
  std::vector<Tensor3> cache_;
};

// kokkosZHEEV(A, W) for M column-major N x N matrices stored contiguously
// (common_kokkos.cpp:11-100): eigenvectors overwrite A, eigenvalues ascending.
inline void batchedZHEEV(int M, int N, std::complex<double>* A, double* W) {
  check(phx_zheev_batched(M, N, reinterpret_cast<double*>(A), W, nullptr));
}

}  // namespace phx_adaptor
