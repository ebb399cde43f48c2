// Phoebe's electron-phonon coupling file (see elph_file.hpp).  Host code only.
#include "elph_file.hpp"

#include <algorithm>
#include <fstream>
#include <stdexcept>
#include <vector>

namespace phx {

namespace {
[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

// The reference checks existence first with a plain open and reports the
// missing file by name (interaction_elph_parsing.cpp:629-636).
const std::string& require_file(const std::string& path) {
  std::ifstream f(path);
  if (!f.is_open())
    fail("Required electron-phonon file ***.phoebe.elph.hdf5 not found at " + path);
  return path;
}
}  // namespace

int64_t ElphFile::scalar(const char* name) const {
  const h5::Dataset& d = r_.get(name);
  if (d.count() != 1) fail(std::string("elph file: ") + name + " is not a scalar");
  int64_t v = 0;
  if (d.type == h5::Type::Int) {
    r_.read_int64(d, 0, 1, &v);
  } else {
    double x = 0;
    r_.read_double(d, 0, 1, &x);
    v = int64_t(x);
    if (double(v) != x) fail(std::string("elph file: ") + name + " is not an integer");
  }
  return v;
}

ElphFile::ElphFile(const std::string& path) : r_(require_file(path)) {
  // a file without /fileFormat predates format 2 (interaction_elph_parsing.cpp:651-663)
  h_.fileFormat = r_.has("/fileFormat") ? int(scalar("/fileFormat")) : 1;
  if (h_.fileFormat != 1 && h_.fileFormat != 2)
    fail("elph file: unknown fileFormat " + std::to_string(h_.fileFormat));
  h_.numElectrons = int(scalar("/numElectrons"));
  h_.numSpin = int(scalar("/numSpin"));
  h_.nw = int(scalar("/numElBands"));
  h_.nm = int(scalar("/numPhModes"));
  if (h_.nw < 1 || h_.nm < 1) fail("elph file: bad numElBands / numPhModes");
  h_.nRe = int(r_.get("/elDegeneracies").count());
  h_.nRp = int(r_.get("/phDegeneracies").count());
  if (h_.nRe < 1 || h_.nRp < 1) fail("elph file: empty lattice vector lists");
  if (r_.get("/elBravaisVectors").count() != 3 * uint64_t(h_.nRe))
    fail("elph file: elBravaisVectors and elDegeneracies disagree");
  if (r_.get("/phBravaisVectors").count() != 3 * uint64_t(h_.nRp))
    fail("elph file: phBravaisVectors and phDegeneracies disagree");
  for (int m = 0; m < 2; ++m) {
    const char* nm = m ? "/qMesh" : "/kMesh";
    if (!r_.has(nm)) continue;
    const h5::Dataset& d = r_.get(nm);
    if (d.count() != 3) fail(std::string("elph file: ") + nm + " is not a 3-vector");
    int64_t v[3];
    r_.read_int64(d, 0, 3, v);
    for (int i = 0; i < 3; ++i) (m ? h_.qMesh : h_.kMesh)[i] = int(v[i]);
  }
  if (h_.fileFormat == 1) {
    const h5::Dataset& g = r_.get("/gWannier");
    if (g.type != h5::Type::Complex) fail("elph file: /gWannier is not complex");
    if (g.count() != gw_elements())
      fail("elph file: /gWannier has " + std::to_string(g.count()) + " elements, the header implies " +
           std::to_string(gw_elements()));
  } else {
    for (int e = 0; e < h_.nRe; ++e) {
      const h5::Dataset& g = r_.get("/gWannier_" + std::to_string(e));
      if (g.type != h5::Type::Complex || g.count() != gw_elements_per_re())
        fail("elph file: /gWannier_" + std::to_string(e) + " has the wrong type or size");
    }
  }
}

uint64_t ElphFile::gw_elements() const { return gw_elements_per_re() * uint64_t(h_.nRe); }

void ElphFile::bravais(double* elR, double* elDeg, double* phR, double* phDeg) const {
  if (elR) r_.read_double(r_.get("/elBravaisVectors"), 0, 3 * uint64_t(h_.nRe), elR);
  if (elDeg) r_.read_double(r_.get("/elDegeneracies"), 0, uint64_t(h_.nRe), elDeg);
  if (phR) r_.read_double(r_.get("/phBravaisVectors"), 0, 3 * uint64_t(h_.nRp), phR);
  if (phDeg) r_.read_double(r_.get("/phDegeneracies"), 0, uint64_t(h_.nRp), phDeg);
}

void ElphFile::read_gw(uint64_t first, uint64_t n, double* out) const {
  const uint64_t total = gw_elements();
  if (first > total || n > total - first) fail("elph file: gWannier read out of range");
  if (h_.fileFormat == 1) {
    r_.read_complex(r_.get("/gWannier"), first, n, out);
    return;
  }
  const uint64_t per = gw_elements_per_re();
  while (n) {
    const uint64_t e = first / per, o = first % per;
    const uint64_t k = std::min(n, per - o);
    r_.read_complex(r_.get("/gWannier_" + std::to_string(e)), o, k, out);
    out += 2 * k;
    first += k;
    n -= k;
  }
}

void write_elph_file(const std::string& path, const ElphHeader& h, const double* elR,
                     const double* elDeg, const double* phR, const double* phDeg, const double* gW) {
  if (h.fileFormat != 1 && h.fileFormat != 2) fail("elph file: fileFormat must be 1 or 2");
  if (h.nw < 1 || h.nm < 1 || h.nRe < 1 || h.nRp < 1) fail("elph file: bad sizes");
  if (!elR || !elDeg || !phR || !phDeg || !gW) fail("elph file: null array");
  const int32_t ints[6] = {h.fileFormat, h.numElectrons, h.numSpin, h.nw, h.nm, 0};
  const int32_t km[3] = {h.kMesh[0], h.kMesh[1], h.kMesh[2]};
  const int32_t qm[3] = {h.qMesh[0], h.qMesh[1], h.qMesh[2]};
  using h5::Type;
  std::vector<h5::WriteSet> s;
  // scalars as HighFive writes an int (rank 0), vectors as Eigen (n, 1) and
  // matrices as Eigen (3, n) with the raw column-major buffer
  s.push_back({"fileFormat", {}, Type::Int, &ints[0]});
  s.push_back({"numElectrons", {}, Type::Int, &ints[1]});
  s.push_back({"numSpin", {}, Type::Int, &ints[2]});
  s.push_back({"numElBands", {}, Type::Int, &ints[3]});
  s.push_back({"numPhModes", {}, Type::Int, &ints[4]});
  s.push_back({"kMesh", {3, 1}, Type::Int, km});
  s.push_back({"qMesh", {3, 1}, Type::Int, qm});
  s.push_back({"phBravaisVectors", {3, uint64_t(h.nRp)}, Type::Float, phR});
  s.push_back({"elBravaisVectors", {3, uint64_t(h.nRe)}, Type::Float, elR});
  s.push_back({"phDegeneracies", {uint64_t(h.nRp), 1}, Type::Float, phDeg});
  s.push_back({"elDegeneracies", {uint64_t(h.nRe), 1}, Type::Float, elDeg});
  const uint64_t per = uint64_t(h.nw) * h.nw * h.nm * h.nRp;
  std::vector<std::string> names;
  if (h.fileFormat == 1) {
    s.push_back({"gWannier", {1, per * uint64_t(h.nRe)}, Type::Complex, gW});
  } else {
    names.reserve(size_t(h.nRe));
    for (int e = 0; e < h.nRe; ++e) names.push_back("gWannier_" + std::to_string(e));
    for (int e = 0; e < h.nRe; ++e)
      s.push_back({names[size_t(e)], {per, 1}, Type::Complex, gW + 2 * per * uint64_t(e)});
  }
  h5::write_file(path, s);
}

}  // namespace phx
