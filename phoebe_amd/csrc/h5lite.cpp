// Minimal HDF5 reader and writer (see h5lite.hpp).  Host code only.
#include "h5lite.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <stdexcept>

namespace phx {
namespace h5 {

namespace {

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error("HDF5: " + m); }

// Little-endian cursor over one loaded structure; every read is bounds-checked.
struct Cur {
  const unsigned char* p;
  size_t n;
  size_t i = 0;
  Cur(const unsigned char* p_, size_t n_) : p(p_), n(n_) {}
  explicit Cur(const std::vector<unsigned char>& v) : p(v.data()), n(v.size()) {}
  void need(size_t k) const {
    if (k > n || i > n - k) fail("truncated or corrupt structure");
  }
  uint64_t u(int bytes) {
    need(size_t(bytes));
    uint64_t v = 0;
    for (int b = 0; b < bytes; ++b) v |= uint64_t(p[i + b]) << (8 * b);
    i += size_t(bytes);
    return v;
  }
  void skip(size_t k) {
    need(k);
    i += k;
  }
  bool sig(const char* s) {
    need(4);
    bool ok = std::memcmp(p + i, s, 4) == 0;
    i += 4;
    return ok;
  }
  std::string cstr() {  // NUL-terminated string
    size_t j = i;
    while (j < n && p[j]) ++j;
    if (j >= n) fail("unterminated name");
    std::string s(reinterpret_cast<const char*>(p + i), j - i);
    i = j + 1;
    return s;
  }
};

std::string strip(const std::string& name) {
  size_t k = 0;
  while (k < name.size() && name[k] == '/') ++k;
  return name.substr(k);
}

// Element type as parsed from a datatype message.
struct TypeDesc {
  int cls = -1;
  uint32_t size = 0;
  bool le = true;
  bool isSigned = false;
  bool ieee = false;  // floating point with the IEEE-754 binary32/64 field layout
  std::vector<std::string> names;
  std::vector<uint32_t> offs;
  std::vector<TypeDesc> members;
};

void parse_type(Cur& c, TypeDesc& t) {
  const unsigned b0 = unsigned(c.u(1));
  t.cls = int(b0 & 15u);
  const unsigned ver = b0 >> 4;
  const uint32_t bits = uint32_t(c.u(3));
  t.size = uint32_t(c.u(4));
  if (t.cls == 0) {  // fixed point
    t.le = (bits & 1u) == 0;
    t.isSigned = (bits & 8u) != 0;
    const unsigned off = unsigned(c.u(2)), prec = unsigned(c.u(2));
    if (off != 0 || prec != 8 * t.size || (t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8))
      t.cls = -2;  // padded or odd-width integers are not read
  } else if (t.cls == 1) {  // floating point
    t.le = (bits & 1u) == 0 && (bits & 0x40u) == 0;
    const unsigned off = unsigned(c.u(2)), prec = unsigned(c.u(2));
    const unsigned eloc = unsigned(c.u(1)), esz = unsigned(c.u(1));
    const unsigned mloc = unsigned(c.u(1)), msz = unsigned(c.u(1));
    const uint32_t bias = uint32_t(c.u(4));
    const unsigned sign = (bits >> 8) & 0xffu;
    t.ieee = (t.size == 8 && off == 0 && prec == 64 && eloc == 52 && esz == 11 && mloc == 0 &&
              msz == 52 && bias == 1023 && sign == 63) ||
             (t.size == 4 && off == 0 && prec == 32 && eloc == 23 && esz == 8 && mloc == 0 &&
              msz == 23 && bias == 127 && sign == 31);
  } else if (t.cls == 6) {  // compound
    const unsigned nmemb = bits & 0xffffu;
    for (unsigned m = 0; m < nmemb; ++m) {
      const size_t start = c.i;
      std::string nm = c.cstr();
      uint32_t off = 0;
      if (ver < 3) {
        const size_t used = c.i - start;
        c.skip((8 - used % 8) % 8);  // name padded to a multiple of 8
        off = uint32_t(c.u(4));
        if (ver == 1) c.skip(1 + 3 + 4 + 4 + 16);  // dimensionality, reserved, permutation, reserved, dims
      } else {
        int ob = t.size < 256u ? 1 : t.size < 65536u ? 2 : t.size < 16777216u ? 3 : 4;
        off = uint32_t(c.u(ob));
      }
      TypeDesc mt;
      parse_type(c, mt);
      t.names.push_back(nm);
      t.offs.push_back(off);
      t.members.push_back(mt);
    }
  } else {
    // other classes: the properties are not needed; the message body bounds the parse
    t.cls = -2;
  }
}

bool is_name(const std::string& s, std::initializer_list<const char*> alts) {
  for (const char* a : alts)
    if (s == a) return true;
  return false;
}

}  // namespace

uint64_t Dataset::count() const {
  if (null) return 0;
  uint64_t c = 1;
  for (uint64_t d : dims) c *= d;
  return c;
}

// ---------------------------------------------------------------- reader

Reader::Reader(const std::string& path) : path_(path) {
  fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd_ < 0) fail("cannot open " + path + ": " + std::strerror(errno));
  try {
    struct stat st;
    if (fstat(fd_, &st) != 0) fail("cannot stat " + path);
    fileSize_ = uint64_t(st.st_size);
    uint64_t root = 0;
    parse_superblock(root);
    read_root(root);
  } catch (...) {
    ::close(fd_);
    fd_ = -1;
    throw;
  }
}

Reader::~Reader() {
  if (fd_ >= 0) ::close(fd_);
}

void Reader::pread_exact(uint64_t off, void* dst, uint64_t n) const {
  if (off > fileSize_ || n > fileSize_ - off) fail("address beyond the end of " + path_);
  auto* d = static_cast<unsigned char*>(dst);
  while (n) {
    const size_t k = size_t(std::min<uint64_t>(n, uint64_t(1) << 30));
    const ssize_t r = ::pread(fd_, d, k, off_t(off));
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) fail("read error in " + path_);
    d += r;
    off += uint64_t(r);
    n -= uint64_t(r);
  }
}

std::vector<unsigned char> Reader::load(uint64_t addr, uint64_t n) const {
  std::vector<unsigned char> v{};
  v.resize(size_t(n));
  pread_exact(base_ + addr, v.data(), n);
  return v;
}

uint64_t Reader::undef() const { return so_ >= 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * so_)) - 1; }

void Reader::parse_superblock(uint64_t& root) {
  static const unsigned char kSig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  uint64_t at = 0;
  for (;;) {
    if (at + 8 > fileSize_) fail(path_ + " is not an HDF5 file");
    unsigned char s[8];
    pread_exact(at, s, 8);
    if (std::memcmp(s, kSig, 8) == 0) break;
    at = at ? at * 2 : 512;
  }
  std::vector<unsigned char> b(size_t(std::min<uint64_t>(256, fileSize_ - at)));
  pread_exact(at, b.data(), b.size());
  Cur c(b);
  c.skip(8);
  const unsigned ver = unsigned(c.u(1));
  if (ver <= 1) {
    c.skip(4);  // free-space, root group, reserved, shared header versions
    so_ = int(c.u(1));
    sl_ = int(c.u(1));
    c.skip(1 + 2 + 2 + 4);  // reserved, group leaf K, group internal K, consistency flags
    if (ver == 1) c.skip(4);
    if (so_ < 2 || so_ > 8 || sl_ < 2 || sl_ > 8) fail("unsupported address sizes");
    base_ = c.u(so_);
    c.skip(3 * size_t(so_));  // free-space info, end of file, driver info
    c.skip(size_t(so_));      // root entry: link name offset
    root = c.u(so_);
  } else if (ver == 2 || ver == 3) {
    so_ = int(c.u(1));
    sl_ = int(c.u(1));
    c.skip(1);
    if (so_ < 2 || so_ > 8 || sl_ < 2 || sl_ > 8) fail("unsupported address sizes");
    base_ = c.u(so_);
    c.skip(2 * size_t(so_));  // extension, end of file
    root = c.u(so_);
  } else {
    fail("unsupported superblock version " + std::to_string(ver));
  }
}

std::vector<Reader::Msg> Reader::messages(uint64_t addr) const {
  std::vector<Msg> out;
  std::vector<std::pair<uint64_t, uint64_t>> blocks;
  if (addr >= fileSize_ || base_ > fileSize_ - addr - 1) fail("object header beyond the end of " + path_);
  std::vector<unsigned char> head = load(addr, std::min<uint64_t>(16, fileSize_ - base_ - addr));
  Cur h(head);
  int version = 0;
  bool order = false;  // v2: messages carry a creation order field
  if (head.size() >= 4 && std::memcmp(head.data(), "OHDR", 4) == 0) {
    version = 2;
    // prefix: signature, version, flags, optional times / attribute limits, chunk-0 size
    std::vector<unsigned char> pre = load(addr, std::min<uint64_t>(64, fileSize_ - base_ - addr));
    Cur c(pre);
    c.skip(4);
    if (c.u(1) != 2) fail("unsupported object header version");
    const unsigned flags = unsigned(c.u(1));
    if (flags & 0x20u) c.skip(16);
    if (flags & 0x10u) c.skip(4);
    const uint64_t size0 = c.u(1 << (flags & 3u));
    order = (flags & 0x04u) != 0;
    blocks.emplace_back(addr + c.i, size0 + 4);  // messages, gap and checksum
  } else {
    version = int(h.u(1));
    if (version != 1) fail("unsupported object header version " + std::to_string(version));
    h.skip(1);
    h.skip(2 + 4);  // message count, reference count
    const uint64_t size = h.u(4);
    blocks.emplace_back(addr + 16, size);
  }
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    if (bi > 256) fail("too many object header continuation blocks");
    const uint64_t baddr = blocks[bi].first, blen = blocks[bi].second;
    std::vector<unsigned char> blk = load(baddr, blen);
    Cur c(blk);
    size_t end = blk.size();
    if (version == 2) {
      if (bi > 0) {
        if (!c.sig("OCHK")) fail("bad object header continuation block");
      }
      if (end < 4) fail("short object header block");
      end -= 4;  // checksum
    }
    const size_t hdr = version == 1 ? 8 : (order ? 6 : 4);
    while (c.i + hdr <= end) {
      unsigned type, flags;
      size_t size;
      if (version == 1) {
        type = unsigned(c.u(2));
        size = size_t(c.u(2));
        flags = unsigned(c.u(1));
        c.skip(3);
      } else {
        type = unsigned(c.u(1));
        size = size_t(c.u(2));
        flags = unsigned(c.u(1));
        if (order) c.skip(2);
      }
      if (c.i + size > end) fail("object header message overruns its block");
      Msg m;
      m.type = type;
      m.body.assign(blk.begin() + long(c.i), blk.begin() + long(c.i + size));
      c.skip(size);
      if (type == 0x10) {  // continuation
        Cur b(m.body);
        const uint64_t ca = b.u(so_), cl = b.u(sl_);
        blocks.emplace_back(ca, cl);
        continue;
      }
      if ((flags & 0x02u) && (type == 0x01 || type == 0x03 || type == 0x08))
        fail("shared dataspace/datatype/layout messages are not supported");
      out.push_back(std::move(m));
    }
  }
  return out;
}

void Reader::read_root(uint64_t addr) {
  const std::vector<Msg> ms = messages(addr);
  bool any = false;
  for (const Msg& m : ms) {
    if (m.type == 0x11) {  // symbol table: v1 B-tree + local heap
      Cur c(m.body);
      const uint64_t btree = c.u(so_), heap = c.u(so_);
      std::vector<unsigned char> hh = load(heap, 8 + 2 * uint64_t(sl_) + uint64_t(so_));
      Cur hc(hh);
      if (!hc.sig("HEAP")) fail("bad local heap signature");
      hc.skip(4);
      const uint64_t dsize = hc.u(sl_);
      hc.skip(size_t(sl_));
      const uint64_t daddr = hc.u(so_);
      std::vector<unsigned char> heapData = load(daddr, dsize);
      walk_btree(btree, heapData, 0);
      any = true;
    } else if (m.type == 0x02) {  // link info: dense storage has a fractal heap
      Cur c(m.body);
      c.skip(1);
      const unsigned flags = unsigned(c.u(1));
      if (flags & 1u) c.skip(8);
      const uint64_t fheap = c.u(so_);
      if (fheap != undef()) fail("groups with dense link storage are not supported");
      any = true;
    } else if (m.type == 0x06) {  // link (compact storage)
      Cur c(m.body);
      if (c.u(1) != 1) fail("unsupported link message version");
      const unsigned flags = unsigned(c.u(1));
      unsigned ltype = 0;
      if (flags & 0x08u) ltype = unsigned(c.u(1));
      if (flags & 0x04u) c.skip(8);
      if (flags & 0x10u) c.skip(1);
      const uint64_t nlen = c.u(1 << (flags & 3u));
      c.need(size_t(nlen));
      std::string name(reinterpret_cast<const char*>(c.p + c.i), size_t(nlen));
      c.skip(size_t(nlen));
      if (ltype == 0) add_object(name, c.u(so_));
      any = true;
    }
  }
  if (!any) fail("root object of " + path_ + " is not a group");
}

void Reader::walk_btree(uint64_t addr, const std::vector<unsigned char>& heap, int depth) {
  if (depth > 32) fail("group B-tree too deep");
  std::vector<unsigned char> hd = load(addr, 8 + 2 * uint64_t(so_));
  Cur h(hd);
  if (!h.sig("TREE")) fail("bad B-tree signature");
  if (h.u(1) != 0) fail("group B-tree of the wrong node type");
  const unsigned level = unsigned(h.u(1));
  const unsigned used = unsigned(h.u(2));
  const uint64_t body = addr + 8 + 2 * uint64_t(so_);
  std::vector<unsigned char> kv = load(body, uint64_t(used) * (uint64_t(sl_) + uint64_t(so_)) + uint64_t(sl_));
  Cur c(kv);
  for (unsigned e = 0; e < used; ++e) {
    c.skip(size_t(sl_));  // key
    const uint64_t child = c.u(so_);
    if (level > 0) {
      walk_btree(child, heap, depth + 1);
      continue;
    }
    std::vector<unsigned char> sh = load(child, 8);
    Cur s(sh);
    if (!s.sig("SNOD")) fail("bad symbol table node signature");
    s.skip(2);
    const unsigned nsym = unsigned(s.u(2));
    const uint64_t esz = 2 * uint64_t(so_) + 24;
    std::vector<unsigned char> ents = load(child + 8, uint64_t(nsym) * esz);
    Cur ec(ents);
    for (unsigned k = 0; k < nsym; ++k) {
      const uint64_t nameOff = ec.u(so_);
      const uint64_t oh = ec.u(so_);
      ec.skip(24);
      if (nameOff >= heap.size()) fail("link name outside the local heap");
      Cur nc(heap.data() + nameOff, heap.size() - size_t(nameOff));
      add_object(nc.cstr(), oh);
    }
  }
}

void Reader::add_object(const std::string& name, uint64_t addr) {
  const std::vector<Msg> ms = messages(addr);
  const Msg *space = nullptr, *type = nullptr, *layout = nullptr;
  for (const Msg& m : ms) {
    if (m.type == 0x01) space = &m;
    if (m.type == 0x03) type = &m;
    if (m.type == 0x08) layout = &m;
    if (m.type == 0x07) fail("dataset " + name + ": external storage is not supported");
    if (m.type == 0x0B) {
      Cur c(m.body);
      c.skip(1);
      if (c.u(1) > 0) fail("dataset " + name + ": filtered (compressed) storage is not supported");
    }
  }
  if (!space || !type || !layout) return;  // not a dataset (a nested group, a named type)
  Dataset d;
  d.name = name;
  {
    Cur c(space->body);
    const unsigned ver = unsigned(c.u(1));
    const unsigned rank = unsigned(c.u(1));
    const unsigned flags = unsigned(c.u(1));
    if (ver == 1) {
      c.skip(5);
    } else if (ver == 2) {
      const unsigned st = unsigned(c.u(1));
      d.null = st == 2;
    } else {
      fail("dataset " + name + ": unsupported dataspace version");
    }
    (void)flags;
    for (unsigned r = 0; r < rank; ++r) d.dims.push_back(c.u(sl_));
  }
  {
    Cur c(type->body);
    TypeDesc t;
    parse_type(c, t);
    d.elemSize = t.size;
    if (t.cls == 0 && t.le) {
      d.type = Type::Int;
      d.isSigned = t.isSigned;
    } else if (t.cls == 1 && t.le && t.ieee) {
      d.type = Type::Float;
    } else if (t.cls == 6 && t.members.size() == 2) {
      const TypeDesc &a = t.members[0], &b = t.members[1];
      if (a.cls == 1 && b.cls == 1 && a.ieee && b.ieee && a.le && b.le && a.size == b.size) {
        int re = 0;
        if (is_name(t.names[1], {"r", "re", "real"}) && is_name(t.names[0], {"i", "im", "imag"})) re = 1;
        d.type = Type::Complex;
        d.partSize = a.size;
        d.reOff = t.offs[size_t(re)];
        d.imOff = t.offs[size_t(1 - re)];
        if (d.reOff + d.partSize > d.elemSize || d.imOff + d.partSize > d.elemSize)
          fail("dataset " + name + ": compound members outside the element");
      }
    }
  }
  {
    Cur c(layout->body);
    const unsigned ver = unsigned(c.u(1));
    int cls = -1;
    if (ver == 1 || ver == 2) {
      const unsigned nd = unsigned(c.u(1));
      cls = int(c.u(1));
      c.skip(5);
      if (cls == 1) {
        d.address = c.u(so_);
        c.skip(size_t(nd) * 4);
      } else if (cls == 0) {
        c.skip(size_t(nd) * 4);
        const uint64_t sz = c.u(4);
        c.need(size_t(sz));
        d.compact.assign(c.p + c.i, c.p + c.i + sz);
      }
    } else if (ver == 3 || ver == 4) {
      cls = int(c.u(1));
      if (cls == 0) {
        const uint64_t sz = c.u(2);
        c.need(size_t(sz));
        d.compact.assign(c.p + c.i, c.p + c.i + sz);
      } else if (cls == 1) {
        d.address = c.u(so_);
        (void)c.u(sl_);
      }
    } else {
      fail("dataset " + name + ": unsupported layout version");
    }
    if (cls != 0 && cls != 1) fail("dataset " + name + ": chunked or virtual storage is not supported");
    d.layout = cls;
    d.allocated = cls == 0 || d.address != undef();
    const uint64_t need = d.count() * d.elemSize;
    if (cls == 0 && d.compact.size() < need) fail("dataset " + name + ": compact data too short");
    if (cls == 1 && d.allocated && (d.address > fileSize_ || need > fileSize_ - base_ - d.address))
      fail("dataset " + name + ": raw data beyond the end of the file");
  }
  sets_[name] = std::move(d);
}

bool Reader::has(const std::string& name) const { return sets_.count(strip(name)) > 0; }

const Dataset& Reader::get(const std::string& name) const {
  auto it = sets_.find(strip(name));
  if (it == sets_.end()) fail("dataset " + name + " not found in " + path_);
  return it->second;
}

std::vector<std::string> Reader::names() const {
  std::vector<std::string> v;
  for (const auto& kv : sets_) v.push_back(kv.first);
  return v;
}

void Reader::raw_read(const Dataset& d, uint64_t byteOff, uint64_t n, void* dst) const {
  const uint64_t total = d.count() * d.elemSize;
  if (byteOff > total || n > total - byteOff) fail("read beyond dataset " + d.name);
  if (d.layout == 0) {
    std::memcpy(dst, d.compact.data() + byteOff, size_t(n));
  } else if (!d.allocated) {
    std::memset(dst, 0, size_t(n));
  } else {
    pread_exact(base_ + d.address + byteOff, dst, n);
  }
}

void Reader::read_complex(const Dataset& d, uint64_t first, uint64_t n, double* out) const {
  if (d.type != Type::Complex) fail("dataset " + d.name + " is not complex");
  if (d.elemSize == 16 && d.partSize == 8 && d.reOff == 0 && d.imOff == 8) {
    raw_read(d, first * 16, n * 16, out);
    return;
  }
  const uint64_t step = 1 << 16;
  std::vector<unsigned char> tmp(size_t(step * d.elemSize));
  for (uint64_t o = 0; o < n; o += step) {
    const uint64_t k = std::min(step, n - o);
    raw_read(d, (first + o) * d.elemSize, k * d.elemSize, tmp.data());
    for (uint64_t e = 0; e < k; ++e) {
      const unsigned char* el = tmp.data() + e * d.elemSize;
      double re, im;
      if (d.partSize == 8) {
        std::memcpy(&re, el + d.reOff, 8);
        std::memcpy(&im, el + d.imOff, 8);
      } else {
        float a, b;
        std::memcpy(&a, el + d.reOff, 4);
        std::memcpy(&b, el + d.imOff, 4);
        re = a;
        im = b;
      }
      out[2 * (o + e)] = re;
      out[2 * (o + e) + 1] = im;
    }
  }
}

void Reader::read_double(const Dataset& d, uint64_t first, uint64_t n, double* out) const {
  if (d.type == Type::Float && d.elemSize == 8) {
    raw_read(d, first * 8, n * 8, out);
    return;
  }
  if (d.type == Type::Float && d.elemSize == 4) {
    std::vector<float> t(static_cast<size_t>(n));
    raw_read(d, first * 4, n * 4, t.data());
    for (uint64_t i = 0; i < n; ++i) out[i] = t[size_t(i)];
    return;
  }
  if (d.type == Type::Int) {
    std::vector<int64_t> t(static_cast<size_t>(n));
    read_int64(d, first, n, t.data());
    for (uint64_t i = 0; i < n; ++i) out[i] = double(t[size_t(i)]);
    return;
  }
  fail("dataset " + d.name + " is not real-valued");
}

void Reader::read_int64(const Dataset& d, uint64_t first, uint64_t n, int64_t* out) const {
  if (d.type != Type::Int) fail("dataset " + d.name + " is not an integer dataset");
  const unsigned s = d.elemSize;
  std::vector<unsigned char> t(static_cast<size_t>(n * s));
  raw_read(d, first * s, n * s, t.data());
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t v = 0;
    for (unsigned b = 0; b < s; ++b) v |= uint64_t(t[size_t(i * s + b)]) << (8 * b);
    if (d.isSigned && s < 8 && (v >> (8 * s - 1)) & 1u) v |= ~uint64_t(0) << (8 * s);
    if (!d.isSigned && s == 8 && (v >> 63)) fail("dataset " + d.name + ": value out of range");
    out[i] = int64_t(v);
  }
}

// ---------------------------------------------------------------- writer

namespace {

struct Buf {
  std::vector<unsigned char> b;
  void u(uint64_t v, int bytes) {
    for (int i = 0; i < bytes; ++i) b.push_back(static_cast<unsigned char>(v >> (8 * i)));
  }
  void raw(const void* p, size_t n) {
    auto* c = static_cast<const unsigned char*>(p);
    b.insert(b.end(), c, c + n);
  }
  void pad8() {
    while (b.size() % 8) b.push_back(0);
  }
  size_t size() const { return b.size(); }
};

void put_at(std::vector<unsigned char>& dst, uint64_t at, const Buf& src) {
  if (at + src.size() > dst.size()) fail("internal layout error");
  std::memcpy(dst.data() + at, src.b.data(), src.size());
}

void float64_type(Buf& t) {
  t.u(0x11, 1);  // version 1, class 1 (floating point)
  t.u(0x20, 1);  // little endian, implied mantissa msb
  t.u(63, 1);    // sign bit position
  t.u(0, 1);
  t.u(8, 4);
  t.u(0, 2);
  t.u(64, 2);
  t.u(52, 1);
  t.u(11, 1);
  t.u(0, 1);
  t.u(52, 1);
  t.u(1023, 4);
}

Buf datatype_body(Type ty) {
  Buf t;
  if (ty == Type::Int) {
    t.u(0x10, 1);  // version 1, class 0 (fixed point)
    t.u(0x08, 1);  // little endian, signed
    t.u(0, 2);
    t.u(4, 4);
    t.u(0, 2);
    t.u(32, 2);
  } else if (ty == Type::Float) {
    float64_type(t);
  } else {
    t.u(0x16, 1);  // version 1, class 6 (compound)
    t.u(2, 2);     // two members
    t.u(0, 1);
    t.u(16, 4);
    const char* nm[2] = {"r", "i"};
    for (int m = 0; m < 2; ++m) {
      Buf n;
      n.raw(nm[m], std::strlen(nm[m]) + 1);
      n.pad8();
      t.raw(n.b.data(), n.size());
      t.u(uint64_t(8 * m), 4);  // byte offset
      t.u(0, 4);                // dimensionality + reserved
      t.u(0, 4);                // dimension permutation
      t.u(0, 4);                // reserved
      t.u(0, 16);               // dimension sizes
      float64_type(t);
    }
  }
  t.pad8();
  return t;
}

uint32_t elem_bytes(Type ty) { return ty == Type::Int ? 4u : ty == Type::Float ? 8u : 16u; }

void message(Buf& oh, unsigned type, unsigned flags, const Buf& body) {
  oh.u(type, 2);
  oh.u(body.size(), 2);
  oh.u(flags, 1);
  oh.u(0, 3);
  oh.raw(body.b.data(), body.size());
}

void pwrite_all(int fd, const void* p, uint64_t n, uint64_t off, const std::string& path) {
  auto* c = static_cast<const unsigned char*>(p);
  while (n) {
    const size_t k = size_t(std::min<uint64_t>(n, uint64_t(1) << 30));
    const ssize_t r = ::pwrite(fd, c, k, off_t(off));
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) fail("write error in " + path + ": " + std::strerror(errno));
    c += r;
    off += uint64_t(r);
    n -= uint64_t(r);
  }
}

}  // namespace

void write_file(const std::string& path, const std::vector<WriteSet>& in) {
  if (in.empty()) fail("nothing to write");
  std::vector<WriteSet> sets = in;
  for (auto& s : sets) {
    s.name = strip(s.name);
    if (s.name.empty() || s.name.find('/') != std::string::npos) fail("bad dataset name '" + s.name + "'");
    if (s.type == Type::Other) fail("dataset " + s.name + ": unsupported element type");
  }
  // symbol table nodes keep their entries sorted by name (byte order)
  std::sort(sets.begin(), sets.end(), [](const WriteSet& a, const WriteSet& b) {
    return std::lexicographical_compare(
        a.name.begin(), a.name.end(), b.name.begin(), b.name.end(),
        [](char x, char y) { return static_cast<unsigned char>(x) < static_cast<unsigned char>(y); });
  });
  for (size_t i = 1; i < sets.size(); ++i)
    if (sets[i].name == sets[i - 1].name) fail("duplicate dataset " + sets[i].name);

  const uint64_t n = sets.size();
  const unsigned leafK = 4, perNode = 2 * leafK;
  const uint64_t nSnod = (n + perNode - 1) / perNode;
  const unsigned intK = unsigned(std::max<uint64_t>(16, (nSnod + 1) / 2));
  if (intK > 32767) fail("too many datasets");
  const uint64_t UNDEF = ~uint64_t(0);

  // local heap: "" at offset 0, then the names, each padded to 8 bytes
  Buf heap;
  heap.u(0, 8);
  std::vector<uint64_t> nameOff(n);
  for (uint64_t i = 0; i < n; ++i) {
    nameOff[i] = heap.size();
    heap.raw(sets[i].name.c_str(), sets[i].name.size() + 1);
    heap.pad8();
  }

  // object headers (version 1): messages, padded with a NIL message to 256 bytes
  std::vector<Buf> ohs(n);
  std::vector<uint64_t> layoutAt(n), rawBytes(n);
  for (uint64_t i = 0; i < n; ++i) {
    const WriteSet& s = sets[i];
    uint64_t count = 1;
    for (uint64_t d : s.dims) count *= d;
    rawBytes[i] = count * elem_bytes(s.type);
    Buf space;
    space.u(1, 1);
    space.u(s.dims.size(), 1);
    space.u(s.dims.empty() ? 0 : 1, 1);  // maximum dimensions present
    space.u(0, 5);
    for (uint64_t d : s.dims) space.u(d, 8);
    for (uint64_t d : s.dims) space.u(d, 8);
    Buf fill;
    fill.u(2, 1);  // version 2
    fill.u(2, 1);  // space allocated late
    fill.u(2, 1);  // fill written if set
    fill.u(1, 1);  // defined, size 0
    fill.u(0, 4);
    Buf layout;
    layout.u(3, 1);
    layout.u(1, 1);      // contiguous
    layout.u(0, 8);      // address, patched below
    layout.u(rawBytes[i], 8);
    layout.pad8();
    Buf msgs;
    message(msgs, 0x01, 0, space);
    message(msgs, 0x03, 1, datatype_body(s.type));
    message(msgs, 0x05, 1, fill);
    layoutAt[i] = msgs.size() + 8 + 2;  // body offset + version/class bytes, relative to the messages
    message(msgs, 0x08, 0, layout);
    unsigned nmsg = 4;
    if (msgs.size() < 256) {
      Buf nil;
      nil.u(0, 256 - msgs.size() - 8);
      message(msgs, 0x00, 0, nil);
      ++nmsg;
    }
    Buf& oh = ohs[i];
    oh.u(1, 1);
    oh.u(0, 1);
    oh.u(nmsg, 2);
    oh.u(1, 4);  // reference count
    oh.u(msgs.size(), 4);
    oh.u(0, 4);  // alignment
    oh.raw(msgs.b.data(), msgs.size());
  }

  // addresses
  const uint64_t aRoot = 96;
  const uint64_t rootSize = 16 + 8 + 16;
  const uint64_t aTree = aRoot + rootSize;
  const uint64_t treeSize = 24 + (2 * uint64_t(intK) + 1) * 8 + 2 * uint64_t(intK) * 8;
  const uint64_t aHeap = aTree + treeSize;
  const uint64_t aHeapData = aHeap + 32;
  const uint64_t aSnod = aHeapData + heap.size();
  const uint64_t snodSize = 8 + uint64_t(perNode) * 40;
  uint64_t at = aSnod + nSnod * snodSize;
  std::vector<uint64_t> aOh(n), aRaw(n);
  for (uint64_t i = 0; i < n; ++i) {
    aOh[i] = at;
    at += ohs[i].size();
  }
  const uint64_t metaEnd = at;
  for (uint64_t i = 0; i < n; ++i) {
    at = (at + 7) / 8 * 8;
    aRaw[i] = rawBytes[i] ? at : UNDEF;
    at += rawBytes[i];
  }
  const uint64_t eof = std::max(at, metaEnd);
  for (uint64_t i = 0; i < n; ++i) {
    Buf a;
    a.u(aRaw[i], 8);
    std::memcpy(ohs[i].b.data() + 16 + layoutAt[i], a.b.data(), 8);
  }

  std::vector<unsigned char> meta(size_t(metaEnd), 0);
  {
    Buf sb;
    const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    sb.raw(sig, 8);
    sb.u(0, 4);  // superblock, free-space, root group, reserved
    sb.u(0, 1);  // shared header message format
    sb.u(8, 1);
    sb.u(8, 1);
    sb.u(0, 1);
    sb.u(leafK, 2);
    sb.u(intK, 2);
    sb.u(0, 4);
    sb.u(0, 8);      // base address
    sb.u(UNDEF, 8);  // free-space info
    sb.u(eof, 8);
    sb.u(UNDEF, 8);  // driver info
    sb.u(0, 8);      // root entry: link name offset
    sb.u(aRoot, 8);
    sb.u(1, 4);      // cache type 1: scratch holds the B-tree and heap addresses
    sb.u(0, 4);
    sb.u(aTree, 8);
    sb.u(aHeap, 8);
    put_at(meta, 0, sb);
  }
  {
    Buf r;
    r.u(1, 1);
    r.u(0, 1);
    r.u(1, 2);
    r.u(1, 4);
    r.u(24, 4);
    r.u(0, 4);
    Buf st;
    st.u(aTree, 8);
    st.u(aHeap, 8);
    message(r, 0x11, 0, st);
    put_at(meta, aRoot, r);
  }
  {
    Buf t;
    t.raw("TREE", 4);
    t.u(0, 1);  // group node
    t.u(0, 1);  // leaf level
    t.u(nSnod, 2);
    t.u(UNDEF, 8);
    t.u(UNDEF, 8);
    t.u(0, 8);  // key 0: ""
    for (uint64_t j = 0; j < nSnod; ++j) {
      const uint64_t last = std::min(n, (j + 1) * perNode) - 1;
      t.u(aSnod + j * snodSize, 8);
      t.u(nameOff[last], 8);  // key j+1: the largest name in child j
    }
    put_at(meta, aTree, t);
  }
  {
    Buf h;
    h.raw("HEAP", 4);
    h.u(0, 4);
    h.u(heap.size(), 8);
    h.u(1, 8);  // no free block
    h.u(aHeapData, 8);
    put_at(meta, aHeap, h);
    put_at(meta, aHeapData, heap);
  }
  for (uint64_t j = 0; j < nSnod; ++j) {
    Buf s;
    s.raw("SNOD", 4);
    s.u(1, 1);
    s.u(0, 1);
    const uint64_t b = j * perNode, e = std::min(n, b + perNode);
    s.u(e - b, 2);
    for (uint64_t i = b; i < e; ++i) {
      s.u(nameOff[i], 8);
      s.u(aOh[i], 8);
      s.u(0, 4);
      s.u(0, 4);
      s.u(0, 16);
    }
    put_at(meta, aSnod + j * snodSize, s);
  }
  for (uint64_t i = 0; i < n; ++i) put_at(meta, aOh[i], ohs[i]);

  ::unlink(path.c_str());
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0) fail("cannot create " + path + ": " + std::strerror(errno));
  try {
    pwrite_all(fd, meta.data(), meta.size(), 0, path);
    for (uint64_t i = 0; i < n; ++i) {
      if (!rawBytes[i]) continue;
      if (!sets[i].data) fail("dataset " + sets[i].name + ": no data");
      pwrite_all(fd, sets[i].data, rawBytes[i], aRaw[i], path);
    }
    if (::ftruncate(fd, off_t(eof)) != 0) fail("cannot size " + path);
  } catch (...) {
    ::close(fd);
    throw;
  }
  if (::close(fd) != 0) fail("cannot close " + path);
}

}  // namespace h5
}  // namespace phx
