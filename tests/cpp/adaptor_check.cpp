// Compile check of the C++ adaptor with minimal Eigen-like types (the same
// column-major storage the reference's Eigen types use).
#include <array>
#include <complex>
#include <vector>

#include "../../phoebe_amd/csrc/phx_adaptor.hpp"

struct MatrixXcd {
  int r = 0, c = 0;
  std::vector<std::complex<double>> v;
  MatrixXcd(int r_, int c_) : r(r_), c(c_), v(size_t(r_) * c_) {}
  int rows() const { return r; }
  int cols() const { return c; }
  size_t size() const { return v.size(); }
  const std::complex<double>* data() const { return v.data(); }
};
struct Vector3d {
  std::array<double, 3> v{};
  const double* data() const { return v.data(); }
};
struct VectorXcd {
  std::vector<std::complex<double>> v;
  int size() const { return int(v.size()); }
  const std::complex<double>* data() const { return v.data(); }
};
struct Tensor3 {
  std::vector<double> v;
  Tensor3(int a, int b, int c) : v(size_t(a) * b * c) {}
  double* data() { return v.data(); }
};

using Inter = phx_adaptor::InteractionElPhWan<MatrixXcd, Vector3d, Tensor3>;

int main(int argc, char**) {
  if (argc > 100) {  // never executed: instantiate every member
    std::vector<std::complex<double>> gW(1);
    double R[3] = {0, 0, 0}, d[1] = {1};
    Inter e(gW.data(), 1, 3, 1, 1, R, d, R, d);
    MatrixXcd U(1, 1);
    Vector3d k;
    e.cacheElPh(U, k);
    std::vector<MatrixXcd> U2s{U}, e3s{MatrixXcd(3, 3)};
    std::vector<Vector3d> qs{k};
    std::vector<VectorXcd> polar;
    e.calcCouplingSquared(U, U2s, e3s, qs, polar);
    (void)e.getCouplingSquared(0);
    (void)e.getCouplingDimensions();
    (void)e.estimateNumBatches(1, 1);
    phx_adaptor::batchedZHEEV(1, 1, gW.data(), d);
  }
  return 0;
}
