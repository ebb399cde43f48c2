// Readers of the harmonic inputs (see harmonic_io.hpp).  Host code only.
#include "harmonic_io.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace phx {

namespace {

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

std::vector<std::string> words(const std::string& s) {
  std::vector<std::string> w;
  std::istringstream is(s);
  for (std::string t; is >> t;) w.push_back(t);
  return w;
}

struct LineReader {
  std::ifstream in;
  std::string path;
  long line = 0;
  explicit LineReader(const std::string& p) : in(p), path(p) {}
  std::string next() {
    std::string s;
    if (!std::getline(in, s)) fail(path + ": unexpected end of file after line " + std::to_string(line));
    ++line;
    return s;
  }
  std::string next_nonblank() {
    for (;;) {
      std::string s = next();
      if (s.find_first_not_of(" \t\r") != std::string::npos) return s;
    }
  }
  std::vector<std::string> fields(size_t need) {
    std::vector<std::string> w = words(next());
    if (w.size() < need)
      fail(path + ":" + std::to_string(line) + ": expected " + std::to_string(need) + " fields");
    return w;
  }
};

double num(const std::string& s, const LineReader& r) {
  try {
    size_t k = 0;
    // Fortran exponents (1.0D-3) are read as C ones
    std::string t = s;
    for (char& c : t)
      if (c == 'D' || c == 'd') c = 'E';
    const double v = std::stod(t, &k);
    if (k != t.size()) throw std::invalid_argument(s);
    return v;
  } catch (const std::exception&) {
    fail(r.path + ":" + std::to_string(r.line) + ": cannot read a number from '" + s + "'");
  }
}

int inum(const std::string& s, const LineReader& r) {
  const double v = num(s, r);
  if (v != std::floor(v)) fail(r.path + ":" + std::to_string(r.line) + ": expected an integer");
  return int(v);
}

// Cell vectors a1, a2, a3 (columns of A) from QE's ibrav and celldm,
// following the lattice definitions of Quantum ESPRESSO's latgen as
// restated in qe_input_parser.cpp:18-383 (the common lattices).
void qe_lattice(int ibrav, double* celldm, double* A /* col-major, in: ibrav 0 cell */) {
  double a1[3] = {A[0], A[1], A[2]}, a2[3] = {A[3], A[4], A[5]}, a3[3] = {A[6], A[7], A[8]};
  auto nrm = [](const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
  if (ibrav == 0) {
    if (nrm(a1) == 0. || nrm(a2) == 0. || nrm(a3) == 0.) fail("QE fc: wrong cell for ibrav = 0");
    if (celldm[0] != 0.) {
      for (int i = 0; i < 9; ++i) A[i] *= celldm[0];
    } else {
      celldm[0] = nrm(a1);
    }
    return;
  }
  if (celldm[0] <= 0.) fail("QE fc: wrong celldm(1)");
  const double a = celldm[0];
  for (int i = 0; i < 3; ++i) a1[i] = a2[i] = a3[i] = 0.;
  if (ibrav == 1) {
    a1[0] = a;
    a2[1] = a;
    a3[2] = a;
  } else if (ibrav == 2) {
    const double t = a / 2.;
    a1[0] = -t;
    a1[2] = t;
    a2[1] = t;
    a2[2] = t;
    a3[0] = -t;
    a3[1] = t;
  } else if (ibrav == 3 || ibrav == -3) {
    const double t = a / 2.;
    for (int i = 0; i < 3; ++i) a1[i] = a2[i] = a3[i] = t;
    if (ibrav < 0) {
      a1[0] = -a1[0];
      a2[1] = -a2[1];
      a3[2] = -a3[2];
    } else {
      a2[0] = -a2[0];
      a3[0] = -a3[0];
      a3[1] = -a3[1];
    }
  } else if (ibrav == 4) {
    if (celldm[2] <= 0.) fail("QE fc: wrong celldm(3)");
    const double sr3 = 1.732050807569;
    a1[0] = a;
    a2[0] = -a / 2.;
    a2[1] = a * sr3 / 2.;
    a3[2] = a * celldm[2];
  } else if (ibrav == 6) {
    if (celldm[2] <= 0.) fail("QE fc: wrong celldm(3)");
    a1[0] = a;
    a2[1] = a;
    a3[2] = a * celldm[2];
  } else if (ibrav == 7) {
    if (celldm[2] <= 0.) fail("QE fc: wrong celldm(3)");
    a2[0] = a / 2.;
    a2[1] = a2[0];
    a2[2] = celldm[2] * a / 2.;
    a1[0] = a2[0];
    a1[1] = -a2[0];
    a1[2] = a2[2];
    a3[0] = -a2[0];
    a3[1] = -a2[0];
    a3[2] = a2[2];
  } else if (ibrav == 8) {
    if (celldm[1] <= 0. || celldm[2] <= 0.) fail("QE fc: wrong celldm(2) or celldm(3)");
    a1[0] = a;
    a2[1] = a * celldm[1];
    a3[2] = a * celldm[2];
  } else {
    fail("QE fc: ibrav = " + std::to_string(ibrav) +
         " is not supported (supported: 0, 1, 2, 3, -3, 4, 6, 7, 8)");
  }
  for (int i = 0; i < 3; ++i) {
    A[i] = a1[i];
    A[3 + i] = a2[i];
    A[6 + i] = a3[i];
  }
}

void reciprocal(const double* A, double* B, double& vol) {
  // B = 2 pi A^-T (column-major); volume = |det A|
  const double det = A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
                     A[6] * (A[1] * A[5] - A[4] * A[2]);
  if (det == 0.) fail("singular cell");
  vol = std::fabs(det);
  // inverse of the column-major A: inv(r, c)
  double inv[9];
  inv[0] = (A[4] * A[8] - A[7] * A[5]) / det;
  inv[3] = -(A[3] * A[8] - A[6] * A[5]) / det;
  inv[6] = (A[3] * A[7] - A[6] * A[4]) / det;
  inv[1] = -(A[1] * A[8] - A[7] * A[2]) / det;
  inv[4] = (A[0] * A[8] - A[6] * A[2]) / det;
  inv[7] = -(A[0] * A[7] - A[6] * A[1]) / det;
  inv[2] = (A[1] * A[5] - A[4] * A[2]) / det;
  inv[5] = -(A[0] * A[5] - A[3] * A[2]) / det;
  inv[8] = (A[0] * A[4] - A[3] * A[1]) / det;
  // B(r, c) = 2 pi inv(c, r)
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B[r + 3 * c] = units::twoPi * inv[c + 3 * r];
}

// Weight of r in the Wigner-Seitz cell of the supercell whose neighbour
// vectors are rws (phonon_h0.cpp:505-535): 1 inside, 0 outside, 1/N on the
// border shared with N - 1 equivalent points.
double ws_weight(const double* r, const std::vector<double>& rws) {
  int eq = 1;
  for (size_t k = 0; k < rws.size() / 3; ++k) {
    const double* w = &rws[3 * k];
    const double rrt = r[0] * w[0] + r[1] * w[1] + r[2] * w[2];
    const double ck = rrt - (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) / 2.;
    if (ck > 1.0e-5) return 0.;
    if (std::fabs(ck) <= 1.0e-5) eq += 1;
  }
  return 1. / double(eq);
}

}  // namespace

PhononModel read_qe_fc(const std::string& path, const std::string& sumRuleIn) {
  std::string sumRule = sumRuleIn;
  std::transform(sumRule.begin(), sumRule.end(), sumRule.begin(), ::tolower);
  if (!sumRule.empty() && sumRule != "simple" && sumRule != "crystal")
    fail("invalid Acoustic Sum Rule");  // phonon_h0_sum_rule.cpp:36-38
  if (sumRule == "crystal") fail("the 'crystal' acoustic sum rule is not supported (use 'simple')");

  {
    std::ifstream probe(path);
    if (!probe.is_open()) fail("Dynamical matrix file not found");  // qe_input_parser.cpp:403-405
  }
  LineReader r(path);
  PhononModel m;
  std::vector<std::string> w = r.fields(9);
  const int ntyp = inum(w[0], r), nat = inum(w[1], r), ibrav = inum(w[2], r);
  if (ntyp < 1 || nat < 1) fail(path + ": bad species / atom counts");
  double celldm[6];
  for (int i = 0; i < 6; ++i) celldm[i] = num(w[3 + i], r);
  if (ibrav == 0) {
    for (int i = 0; i < 3; ++i) {
      std::vector<std::string> c = r.fields(3);
      for (int j = 0; j < 3; ++j) m.cell[j + 3 * i] = num(c[size_t(j)], r);
    }
  }
  qe_lattice(ibrav, celldm, m.cell);
  reciprocal(m.cell, m.recip, m.volume);

  for (int t = 0; t < ntyp; ++t) {
    const std::string s = r.next();
    // index 'name' mass
    const size_t q1 = s.find('\''), q2 = s.find('\'', q1 == std::string::npos ? 0 : q1 + 1);
    if (q1 == std::string::npos || q2 == std::string::npos) fail(path + ": bad species line");
    std::string name = s.substr(q1 + 1, q2 - q1 - 1);
    name.erase(std::remove_if(name.begin(), name.end(), ::isspace), name.end());
    std::vector<std::string> rest = words(s.substr(q2 + 1));
    if (rest.empty()) fail(path + ": species mass missing");
    m.speciesNames.push_back(name);
    m.speciesMasses.push_back(num(rest[0], r));
  }
  m.nat = nat;
  m.nm = 3 * nat;
  for (int a = 0; a < nat; ++a) {
    std::vector<std::string> f = r.fields(5);
    const int sp = inum(f[1], r) - 1;
    if (sp < 0 || sp >= ntyp) fail(path + ": atom species out of range");
    m.species.push_back(sp);
    // positions are written in units of alat
    for (int i = 0; i < 3; ++i) m.positions.push_back(celldm[0] * num(f[size_t(2 + i)], r));
  }
  for (int a = 0; a < nat; ++a)
    for (int i = 0; i < 3; ++i) m.masses.push_back(m.speciesMasses[size_t(m.species[size_t(a)])]);

  {
    std::string s = r.next();
    s.erase(std::remove_if(s.begin(), s.end(), ::isspace), s.end());
    m.born.assign(size_t(nat) * 9, 0.);
    if (s == "T") {
      for (int i = 0; i < 3; ++i) {
        std::vector<std::string> f = r.fields(3);
        for (int j = 0; j < 3; ++j) m.dielectric[3 * i + j] = num(f[size_t(j)], r);
      }
      for (int a = 0; a < nat; ++a) {
        r.next();  // atom index
        for (int i = 0; i < 3; ++i) {
          std::vector<std::string> f = r.fields(3);
          for (int j = 0; j < 3; ++j) m.born[size_t(a) * 9 + 3 * i + j] = num(f[size_t(j)], r);
        }
      }
    }
    double epsSum = 0;
    for (double x : m.dielectric) epsSum += x;
    m.hasDielectric = epsSum > 0.;  // phonon_h0.cpp:35-37
  }
  {
    std::vector<std::string> f = r.fields(3);
    for (int i = 0; i < 3; ++i) m.qGrid[i] = inum(f[size_t(i)], r);
    if (m.qGrid[0] <= 0 || m.qGrid[1] <= 0 || m.qGrid[2] <= 0) fail("qCoarseGrid smaller than zero");
  }
  const int g0 = m.qGrid[0], g1 = m.qGrid[1], g2 = m.qGrid[2];
  const size_t ncell = size_t(g0) * g1 * g2;
  // fc(i, j, n1, n2, n3, na, nb) at [(((((i*3 + j)*nat + na)*nat + nb)*g2 + n3)*g1 + n2)*g0 + n1]
  std::vector<double> fc(9 * size_t(nat) * nat * ncell, 0.);
  auto FC = [&](int i, int j, int n1, int n2, int n3, int na, int nb) -> double& {
    return fc[((((size_t(i * 3 + j) * nat + na) * nat + nb) * g2 + n3) * g1 + n2) * g0 + n1];
  };
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int na = 0; na < nat; ++na)
        for (int nb = 0; nb < nat; ++nb) {
          r.next();  // "i j na nb"
          for (int n3 = 0; n3 < g2; ++n3)
            for (int n2 = 0; n2 < g1; ++n2)
              for (int n1 = 0; n1 < g0; ++n1) {
                std::vector<std::string> f = r.fields(4);
                FC(i, j, n1, n2, n3, na, nb) = num(f[3], r);
              }
        }

  // ---- acoustic sum rule (phonon_h0_sum_rule.cpp:44-76)
  if (sumRule == "simple") {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double sum = 0.;
        for (int a = 0; a < nat; ++a) sum += m.born[size_t(a) * 9 + 3 * i + j];
        for (int a = 0; a < nat; ++a) m.born[size_t(a) * 9 + 3 * i + j] -= sum / nat;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int na = 0; na < nat; ++na) {
          double sum = 0.;
          for (int nb = 0; nb < nat; ++nb)
            for (int n1 = 0; n1 < g0; ++n1)
              for (int n2 = 0; n2 < g1; ++n2)
                for (int n3 = 0; n3 < g2; ++n3) sum += FC(i, j, n1, n2, n3, na, nb);
          FC(i, j, 0, 0, 0, na, na) -= sum;
        }
  }

  // ---- Wigner-Seitz weights and the reorder (phonon_h0.cpp:426-503, 672-771)
  const double* A = m.cell;
  double sup[9];
  for (int i = 0; i < 3; ++i) {
    sup[i] = A[i] * g0;
    sup[3 + i] = A[3 + i] * g1;
    sup[6 + i] = A[6 + i] * g2;
  }
  std::vector<double> rws;
  for (int ir = -2; ir <= 2; ++ir)
    for (int jr = -2; jr <= 2; ++jr)
      for (int kr = -2; kr <= 2; ++kr) {
        double v[3];
        for (int i = 0; i < 3; ++i) v[i] = sup[i] * ir + sup[3 + i] * jr + sup[6 + i] * kr;
        if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 1.0e-6) rws.insert(rws.end(), v, v + 3);
      }
  const int b1 = 2 * g0, b2 = 2 * g1, b3 = 2 * g2;
  const int e1 = 2 * b1 + 1, e2 = 2 * b2 + 1, e3 = 2 * b3 + 1;
  // ws(n3, n2, n1, nb, na) at [(((n3*e2 + n2)*e1 + n1)*nat + nb)*nat + na]
  std::vector<double> ws(size_t(e1) * e2 * e3 * nat * nat, 0.);
  auto WS = [&](int n3, int n2, int n1, int nb, int na) -> double& {
    return ws[((size_t(n3 + b3) * e2 + size_t(n2 + b2)) * e1 + size_t(n1 + b1)) * nat * nat + size_t(nb) * nat + na];
  };
  for (int na = 0; na < nat; ++na)
    for (int nb = 0; nb < nat; ++nb) {
      double total = 0.;
      for (int n1 = -b1; n1 <= b1; ++n1)
        for (int n2 = -b2; n2 <= b2; ++n2)
          for (int n3 = -b3; n3 <= b3; ++n3) {
            double rv[3];
            for (int i = 0; i < 3; ++i)
              rv[i] = double(n1) * A[i] + double(n2) * A[3 + i] + double(n3) * A[6 + i] +
                      (m.positions[size_t(na) * 3 + i] - m.positions[size_t(nb) * 3 + i]);
            const double x = ws_weight(rv, rws);
            WS(n3, n2, n1, nb, na) = x;
            total += x;
          }
      if (std::fabs(total - double(ncell)) > 1.0e-8)
        fail("wrong total Wigner-Seitz weight " + std::to_string(total));
    }
  const int nm = m.nm;
  for (int n3 = -b3; n3 <= b3; ++n3)
    for (int n2 = -b2; n2 <= b2; ++n2)
      for (int n1 = -b1; n1 <= b1; ++n1)
        for (int nb = 0; nb < nat; ++nb)
          for (int na = 0; na < nat; ++na) {
            const double wgt = WS(n3, n2, n1, nb, na);
            if (!(wgt > 0.)) continue;
            m.weights.push_back(wgt);
            for (int i = 0; i < 3; ++i)
              m.R.push_back(n1 * A[i] + n2 * A[3 + i] + n3 * A[6 + i]);
            const int m1 = ((n1 % g0) + g0) % g0, m2 = ((n2 % g1) + g1) % g1, m3 = ((n3 % g2) + g2) % g2;
            const size_t base = m.mat2R.size();
            m.mat2R.resize(base + size_t(nm) * nm, 0.);
            for (int j = 0; j < 3; ++j)
              for (int i = 0; i < 3; ++i)
                m.mat2R[base + size_t(na * 3 + i) * nm + size_t(nb * 3 + j)] += FC(i, j, m1, m2, m3, na, nb);
          }
  m.nR = int(m.weights.size());

  // ---- rigid-ion long-range term: G vectors and the constant part (phonon_h0.cpp:53-160)
  m.lrNorm = units::e2 * units::fourPiAsDefined / m.volume;
  m.lrConstant.assign(size_t(nat) * 9, 0.);
  if (m.hasDielectric) {
    const double cutoff = m.gMax * 4.;
    const double* B = m.recip;
    auto coln = [&](int c) {
      return std::sqrt(B[3 * c] * B[3 * c] + B[3 * c + 1] * B[3 * c + 1] + B[3 * c + 2] * B[3 * c + 2]);
    };
    const int nr1x = g0 > 1 ? int(std::sqrt(cutoff) / coln(0)) + 1 : 0;
    const int nr2x = g1 > 1 ? int(std::sqrt(cutoff) / coln(1)) + 1 : 0;
    const int nr3x = g2 > 1 ? int(std::sqrt(cutoff) / coln(2)) + 1 : 0;
    for (int q1 = -nr1x; q1 <= nr1x; ++q1)
      for (int q2 = -nr2x; q2 <= nr2x; ++q2)
        for (int q3 = -nr3x; q3 <= nr3x; ++q3)
          for (int i = 0; i < 3; ++i)
            m.gVectors.push_back(double(q1) * B[i] + double(q2) * B[3 + i] + double(q3) * B[6 + i]);
    m.nG = int(m.gVectors.size() / 3);
    const double* eps = m.dielectric;
    std::vector<double> gZ(size_t(3) * nat), fnAt(size_t(nat) * 3);
    for (int ig = 0; ig < m.nG; ++ig) {
      const double* g = &m.gVectors[size_t(ig) * 3];
      double geg = 0.;
      for (int j = 0; j < 3; ++j) {
        double t = 0.;
        for (int i = 0; i < 3; ++i) t += g[i] * eps[3 * i + j];
        geg += t * g[j];
      }
      if (!(0. < geg && geg < 4. * m.gMax)) continue;
      const double normG = m.lrNorm * std::exp(-geg * 0.25) / geg;
      for (int a = 0; a < nat; ++a)
        for (int i = 0; i < 3; ++i)
          gZ[size_t(i) * nat + a] = g[0] * m.born[size_t(a) * 9 + 0 * 3 + i] +
                                    g[1] * m.born[size_t(a) * 9 + 1 * 3 + i] +
                                    g[2] * m.born[size_t(a) * 9 + 2 * 3 + i];
      std::fill(fnAt.begin(), fnAt.end(), 0.);
      for (int na = 0; na < nat; ++na)
        for (int i = 0; i < 3; ++i)
          for (int nb = 0; nb < nat; ++nb) {
            double arg = 0.;
            for (int k = 0; k < 3; ++k)
              arg += (m.positions[size_t(na) * 3 + k] - m.positions[size_t(nb) * 3 + k]) * g[k];
            fnAt[size_t(na) * 3 + i] += gZ[size_t(i) * nat + nb] * std::cos(arg);
          }
      for (int na = 0; na < nat; ++na)
        for (int j = 0; j < 3; ++j)
          for (int i = 0; i < 3; ++i)
            m.lrConstant[size_t(na) * 9 + 3 * i + j] += -normG * gZ[size_t(i) * nat + na] * fnAt[size_t(na) * 3 + j];
    }
  }
  return m;
}

ElectronModel read_w90_tb(const std::string& path, const double* cell) {
  {
    std::ifstream probe(path);
    if (!probe.is_open()) fail("Wannier H0 file not found");  // qe_input_parser.cpp:905-907
  }
  LineReader r(path);
  ElectronModel e;
  r.next();  // title and date
  double fileCell[9];
  for (int i = 0; i < 3; ++i) {
    std::vector<std::string> f = r.fields(3);
    for (int j = 0; j < 3; ++j) fileCell[j + 3 * i] = num(f[size_t(j)], r) / units::distanceBohrToAng;
  }
  e.nw = inum(r.fields(1)[0], r);
  e.nR = inum(r.fields(1)[0], r);
  if (e.nw < 1 || e.nR < 1) fail(path + ": bad numbers of Wannier functions / vectors");
  const int nw = e.nw, nR = e.nR;
  while (int(e.deg.size()) < nR) {
    for (const std::string& s : words(r.next())) e.deg.push_back(double(inum(s, r)));
  }
  if (int(e.deg.size()) != nR) fail(path + ": degeneracy count mismatch");
  e.Rcrys.resize(size_t(nR) * 3);
  e.h0R.resize(size_t(nR) * nw * nw * 2);
  e.rMat.assign(size_t(3) * nR * nw * nw * 2, 0.);
  for (int iR = 0; iR < nR; ++iR) {
    std::vector<std::string> f = words(r.next_nonblank());
    if (f.size() < 3) fail(path + ":" + std::to_string(r.line) + ": expected a lattice vector");
    for (int i = 0; i < 3; ++i) e.Rcrys[size_t(iR) * 3 + i] = num(f[size_t(i)], r);
    // the reference stores the (i, j) line of the block at h0R(iR, i, j), i slowest
    for (int i = 0; i < nw; ++i)
      for (int j = 0; j < nw; ++j) {
        std::vector<std::string> v = r.fields(4);
        const size_t o = ((size_t(iR) * nw + i) * nw + j) * 2;
        e.h0R[o] = num(v[2], r) / units::energyRyToEv;
        e.h0R[o + 1] = num(v[3], r) / units::energyRyToEv;
      }
  }
  for (int iR = 0; iR < nR; ++iR) {
    r.next_nonblank();
    for (int i = 0; i < nw; ++i)
      for (int j = 0; j < nw; ++j) {
        std::vector<std::string> v = r.fields(8);
        for (int d = 0; d < 3; ++d) {
          const size_t o = (((size_t(d) * nR + iR) * nw + i) * nw + j) * 2;
          e.rMat[o] = num(v[size_t(2 + 2 * d)], r) / units::distanceBohrToAng;
          e.rMat[o + 1] = num(v[size_t(3 + 2 * d)], r) / units::distanceBohrToAng;
        }
      }
  }
  for (int i = 0; i < 9; ++i) e.cell[i] = cell ? cell[i] : fileCell[i];
  // Cartesian vectors: A times the crystal coordinates (qe_input_parser.cpp:1009-1012)
  e.R.resize(size_t(nR) * 3);
  for (int iR = 0; iR < nR; ++iR)
    for (int i = 0; i < 3; ++i) {
      const double* c = &e.Rcrys[size_t(iR) * 3];
      e.R[size_t(iR) * 3 + i] = e.cell[i] * c[0] + e.cell[3 + i] * c[1] + e.cell[6 + i] * c[2];
    }
  return e;
}

}  // namespace phx
