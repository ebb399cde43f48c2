// Header-only C++ adaptor with the call signatures of Phoebe's
// InteractionElPhWan (src/interaction/interaction_elph.h:30-219) and
// kokkosZHEEV (src/common_kokkos.h:49) on top of the C ABI in
// include/phx_elph.h.
//
// Templated on the matrix / tensor types so that it compiles against Eigen
// (Eigen::MatrixXcd, Eigen::Vector3d, Eigen::Tensor<double,3>) inside Phoebe,
// or against any type with the same column-major storage, data() and
// rows()/cols() (or dimension(i)) accessors.  Errors follow the reference's
// Error() behaviour (exceptions.cpp:6-16): print and terminate, unless the
// caller installs its own handler.
#pragma once

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
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
// matrix type (rows(), cols(), data()); Vec3: 3-vector with data();
// Tensor3: rank-3 double tensor with the constructor Tensor3(nb1, nb2, nm)
// and column-major data().
template <class Matrix, class Vec3, class Tensor3>
class InteractionElPhWan {
 public:
  // gW: the reference's Eigen tensor (nw, nw, nm, nRp, nRe) column-major as a
  // flat complex array; elR/phR: (3, nR) column-major Cartesian vectors.
  InteractionElPhWan(const std::complex<double>* gW, int nw, int nm, int nRp, int nRe,
                     const double* elR, const double* elDeg, const double* phR,
                     const double* phDeg, int precision = PHX_FP64)
      : nw_(nw), nm_(nm) {
    check(phx_elph_create(&h_, reinterpret_cast<const double*>(gW), nw, nm, nRp, nRe, elR, elDeg,
                          phR, phDeg, precision));
  }
  InteractionElPhWan(const InteractionElPhWan&) = delete;
  InteractionElPhWan& operator=(const InteractionElPhWan&) = delete;
  ~InteractionElPhWan() {
    if (h_) phx_elph_destroy(h_);
  }

  // interaction_elph.h:161
  void cacheElPh(const Matrix& eigvec1, const Vec3& k1C) {
    nb1_ = int(eigvec1.cols());
    check(phx_elph_cache(h_, reinterpret_cast<const double*>(eigvec1.data()), nb1_, k1C.data()));
  }

  // interaction_elph.h:149-154 (polarData must be empty or all zero)
  template <class VecC>
  void calcCouplingSquared(const Matrix& eigvec1, const std::vector<Matrix>& eigvecs2,
                           const std::vector<Matrix>& eigvecs3, const std::vector<Vec3>& q3Cs,
                           const std::vector<VecC>& polarData) {
    for (const auto& p : polarData)
      for (int i = 0; i < int(p.size()); ++i)
        if (p.data()[i] != std::complex<double>(0.))
          error_handler()("polar correction is not supported by the GPU path");
    const int nq = int(eigvecs2.size());
    const int nb1 = int(eigvec1.cols());
    std::vector<int> nb2s(nq);
    std::vector<std::complex<double>> U2, e3;
    std::vector<double> q3(3 * size_t(nq));
    size_t nout = 0;
    for (int i = 0; i < nq; ++i) {
      nb2s[i] = int(eigvecs2[i].cols());
      U2.insert(U2.end(), eigvecs2[i].data(), eigvecs2[i].data() + eigvecs2[i].size());
      e3.insert(e3.end(), eigvecs3[i].data(), eigvecs3[i].data() + eigvecs3[i].size());
      for (int d = 0; d < 3; ++d) q3[3 * i + d] = q3Cs[i].data()[d];
      nout += size_t(nb1) * nb2s[i] * nm_;
    }
    std::vector<double> out(nout);
    check(phx_elph_coupling_sq(h_, nq, reinterpret_cast<const double*>(U2.data()), nb2s.data(),
                               reinterpret_cast<const double*>(e3.data()), q3.data(), nullptr,
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

  // interaction_elph.h:197, returns (nRe, nRp, nm, nw, nw) as the reference does
  std::vector<int> getCouplingDimensions() {
    std::vector<int> d(5);
    check(phx_elph_get_coupling_dims(h_, d.data()));
    return d;
  }

  // interaction_elph.h:205
  int estimateNumBatches(const int& nk2, const int& nb1) {
    int n = phx_elph_estimate_num_batches(h_, nk2, nb1, 0.0);
    check(n);
    return n;
  }

 private:
  phx_elph* h_ = nullptr;
  int nw_ = 0, nm_ = 0, nb1_ = 0;
  std::vector<Tensor3> cache_;
};

// kokkosZHEEV(A, W) for M column-major N x N matrices stored contiguously
// (common_kokkos.cpp:11-100): eigenvectors overwrite A, eigenvalues ascending.
inline void batchedZHEEV(int M, int N, std::complex<double>* A, double* W) {
  check(phx_zheev_batched(M, N, reinterpret_cast<double*>(A), W, nullptr));
}

}  // namespace phx_adaptor
