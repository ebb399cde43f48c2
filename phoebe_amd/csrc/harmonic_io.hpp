// Readers of the harmonic inputs of the el-ph path, and the set-up work the
// reference does once in its constructors.  Host code only.
//
//   * Quantum ESPRESSO force constants (q2r.x output, prefix.fc):
//     QEParser::parsePhHarmonic (qe_input_parser.cpp:385-562) and the
//     PhononH0 constructor (phonon_h0.cpp:18-235): acoustic sum rule on the
//     force constants and Born charges, Wigner-Seitz reorder of the force
//     constants into (R, weight, block) triples, and the constant part and
//     G-vector list of the rigid-ion long-range term.
//   * Wannier90 tight-binding file (seedname_tb.dat):
//     QEParser::parseElHarmonicWannier (qe_input_parser.cpp:883-1040).
#pragma once

#include <string>
#include <vector>

namespace phx {

struct PhononModel {
  double cell[9] = {};    // direct cell, column-major: column i = lattice vector a_i (Bohr)
  double recip[9] = {};   // 2 pi A^-T, column-major
  double volume = 0;      // Bohr^3
  int nat = 0, nm = 0;
  int qGrid[3] = {0, 0, 0};
  std::vector<double> positions;  // nat x 3, Cartesian Bohr
  std::vector<double> masses;     // nm: the species mass of each mode's atom (Ry units)
  std::vector<int> species;       // nat
  std::vector<double> speciesMasses;
  std::vector<std::string> speciesNames;
  bool hasDielectric = false;
  double dielectric[9] = {};      // row-major (i, j)
  std::vector<double> born;       // nat x 3 x 3, (atom, i, j) after the sum rule
  // Wigner-Seitz form of the force constants (phonon_h0.cpp:672-771)
  int nR = 0;
  std::vector<double> R;          // nR x 3, Cartesian Bohr
  std::vector<double> weights;    // nR
  std::vector<double> mat2R;      // nR x nm x nm, row-major: (R, iAt*3+i, jAt*3+j)
  // rigid-ion long-range term (phonon_h0.cpp:53-160)
  int nG = 0;
  std::vector<double> gVectors;   // nG x 3
  std::vector<double> lrConstant; // nat x 3 x 3: (atom, i, j), added to D(q) on the diagonal blocks
  double lrNorm = 0;              // e2 * fourPi / volume, as the reference defines fourPi
  double gMax = 14.;
};

// sumRule: "" (none), "simple"; "crystal" is rejected as not supported.
PhononModel read_qe_fc(const std::string& path, const std::string& sumRule);

struct ElectronModel {
  double cell[9] = {};        // column-major, Bohr (the file's, or the one passed in)
  int nw = 0, nR = 0;
  std::vector<double> Rcrys;  // nR x 3, integer crystal coordinates as read
  std::vector<double> R;      // nR x 3, Cartesian Bohr
  std::vector<double> deg;    // nR
  std::vector<double> h0R;    // nR x nw x nw complex (interleaved), Ry
  std::vector<double> rMat;   // 3 x nR x nw x nw complex, Bohr (position matrix elements)
};

// cell: column-major direct cell replacing the file's (the reference passes the
// phonon crystal's), or nullptr to use the file's.
ElectronModel read_w90_tb(const std::string& path, const double* cell);

// Constants restated from constants/constants.h:9-60.
namespace units {
constexpr double pi = 3.14159265358979323846;
constexpr double twoPi = 2. * pi;
constexpr double fourPiAsDefined = 2. * pi;  // constants.h:12 defines fourPi as 2 pi
constexpr double e2 = 2.;
constexpr double rydbergSi = 4.359744650e-18 / 2.;
constexpr double electronVoltSi = 1.6021766208e-19;
constexpr double energyRyToEv = rydbergSi / electronVoltSi;
constexpr double bohrRadiusSi = 0.52917721067e-10;
constexpr double distanceBohrToAng = bohrRadiusSi / 1.0e-10;
}  // namespace units

}  // namespace phx
