// Phoebe's electron-phonon coupling file (prefix.phoebe.elph.hdf5) on top of
// the h5lite reader and writer.  Host code only.
//
// Datasets (all in the root group; interaction_elph_parsing.cpp:161-233 and
// qe_to_phoebe_utils.cpp:377-443):
//   /fileFormat       int scalar: 1 (one /gWannier) or 2 (one /gWannier_<irE> per
//                     electronic lattice vector); absent in early files, which are format 1
//   /numElectrons, /numSpin, /numElBands, /numPhModes    int scalars
//   /kMesh, /qMesh    3 ints
//   /elBravaisVectors, /phBravaisVectors   Eigen (3 x nR) doubles, raw column-major
//                     buffer, i.e. R vector r is elements 3r..3r+2 (Bohr)
//   /elDegeneracies, /phDegeneracies       nR doubles
//   /gWannier         format 1: (1, nw*nw*nm*nRp*nRe) complex doubles, the Eigen
//                     tensor (nw, nw, nm, nRp, nRe) in column-major order
//   /gWannier_<irE>   format 2: (nw*nw*nm*nRp, 1) complex doubles for one R_e
#pragma once

#include <cstdint>
#include <string>

#include "h5lite.hpp"

namespace phx {

struct ElphHeader {
  int fileFormat = 1;
  int numElectrons = 0, numSpin = 1;
  int nw = 0, nm = 0;  // numElBands, numPhModes
  int nRe = 0, nRp = 0;
  int kMesh[3] = {0, 0, 0}, qMesh[3] = {0, 0, 0};  // zero when the file has none
};

class ElphFile {
 public:
  // parseHDF5 + parseHeaderHDF5 (interaction_elph_parsing.cpp:626-670, 161-233):
  // format decision and header datasets; the coupling tensor is read on demand.
  explicit ElphFile(const std::string& path);

  const ElphHeader& header() const { return h_; }
  uint64_t gw_elements() const;  // nw^2 nm nRp nRe
  uint64_t gw_elements_per_re() const { return uint64_t(h_.nw) * h_.nw * h_.nm * h_.nRp; }

  // elR, phR: nR x 3 (row r = lattice vector r, Bohr); degeneracies: nR
  void bravais(double* elR, double* elDeg, double* phR, double* phDeg) const;
  // Elements [first, first + n) of the flattened tensor, as interleaved (re, im).
  void read_gw(uint64_t first, uint64_t n, double* out) const;

 private:
  h5::Reader r_;
  ElphHeader h_;
  int64_t scalar(const char* name) const;
};

// writeElPhCouplingHDF5v1 / v2 + writeHeaderHDF5
// (qe_to_phoebe_utils.cpp:446-626, 628-720, 377-443): gW as interleaved
// (re, im) doubles in the tensor order above; elR / phR as nR x 3.
void write_elph_file(const std::string& path, const ElphHeader& h, const double* elR,
                     const double* elDeg, const double* phR, const double* phDeg, const double* gW);

}  // namespace phx
