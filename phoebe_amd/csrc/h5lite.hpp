// Minimal HDF5 reader and writer, without libhdf5, for the files that carry
// Phoebe's electron-phonon coupling in the Wannier representation
// (prefix.phoebe.elph.hdf5): a flat root group of datasets holding integers,
// IEEE doubles or {r, i} pairs of doubles (HighFive's std::complex<double>),
// stored contiguously or compactly, with no filters.
//
// The structures follow the HDF5 File Format Specification, version 3.0
// (The HDF Group): superblock versions 0-3, object header versions 1 and 2,
// root groups indexed by a symbol table (version-1 B-tree of group nodes, a
// local heap and symbol table nodes) or by compact link messages, dataspace
// versions 1 and 2, fixed-point / floating-point / compound datatypes, and
// data layout versions 1-4 (compact and contiguous classes).  The writer emits
// what libhdf5 writes by default for such files: a version-0 superblock, a
// symbol-table root group and version-1 object headers.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace phx {
namespace h5 {

enum class Type { Int, Float, Complex, Other };

struct Dataset {
  std::string name;
  std::vector<uint64_t> dims;  // dataspace extents; empty for a scalar
  bool null = false;           // null dataspace (no elements)
  Type type = Type::Other;
  uint32_t elemSize = 0;       // bytes per element
  bool isSigned = true;        // Int
  uint32_t partSize = 0;       // Complex: bytes of each member (4 or 8)
  uint32_t reOff = 0, imOff = 0;  // Complex: byte offsets of the real / imaginary members
  int layout = -1;             // 0 compact, 1 contiguous
  uint64_t address = 0;        // contiguous: file address of the raw data
  bool allocated = false;      // contiguous: false when the raw data was never written (fill = 0)
  std::vector<unsigned char> compact;
  uint64_t count() const;      // number of elements
};

class Reader {
 public:
  explicit Reader(const std::string& path);
  ~Reader();
  Reader(const Reader&) = delete;
  Reader& operator=(const Reader&) = delete;

  const std::string& path() const { return path_; }
  bool has(const std::string& name) const;            // leading '/' optional
  const Dataset& get(const std::string& name) const;  // throws when absent
  std::vector<std::string> names() const;

  // Elements [first, first + n) converted to the destination type.
  void read_int64(const Dataset& d, uint64_t first, uint64_t n, int64_t* out) const;
  void read_double(const Dataset& d, uint64_t first, uint64_t n, double* out) const;
  // interleaved (re, im) doubles
  void read_complex(const Dataset& d, uint64_t first, uint64_t n, double* out) const;

 private:
  std::string path_;
  int fd_ = -1;
  uint64_t fileSize_ = 0, base_ = 0;
  int so_ = 8, sl_ = 8;  // sizes of offsets and lengths
  std::map<std::string, Dataset> sets_;

  struct Msg {
    unsigned type;
    std::vector<unsigned char> body;
  };
  void pread_exact(uint64_t off, void* dst, uint64_t n) const;
  std::vector<unsigned char> load(uint64_t addr, uint64_t n) const;
  uint64_t undef() const;
  void parse_superblock(uint64_t& rootHeader);
  std::vector<Msg> messages(uint64_t addr) const;
  void read_root(uint64_t addr);
  void walk_btree(uint64_t addr, const std::vector<unsigned char>& heap, int depth);
  void add_object(const std::string& name, uint64_t addr);
  void raw_read(const Dataset& d, uint64_t byteOff, uint64_t n, void* dst) const;
};

// One dataset to write: dims (empty: scalar), element type Int (int32),
// Float (double) or Complex ({r, i} doubles), and its elements in row-major
// order of dims.  Names are root-level ("/name" or "name").
struct WriteSet {
  std::string name;
  std::vector<uint64_t> dims;
  Type type = Type::Float;
  const void* data = nullptr;
};

// Writes a new file (an existing one is removed first).
void write_file(const std::string& path, const std::vector<WriteSet>& sets);

}  // namespace h5
}  // namespace phx
