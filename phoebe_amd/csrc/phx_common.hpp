// Shared helpers for the phx_elph HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace phx {

// Last error message of the calling host thread (returned by phx_last_error).
void set_error(const std::string& msg);
const char* last_error();

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PHX_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::phx::Error(-2, std::string(#call) + ": " + hipGetErrorString(e_) +    \
                                 " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define PHX_REQUIRE(cond, msg)                                   \
  do {                                                           \
    if (!(cond)) throw ::phx::Error(-1, std::string(msg));       \
  } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count) PHX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // grow-only reallocation (contents not preserved)
  void reserve(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Rigid-ion long-range term of the dynamical matrix (device pointers).
struct LongRange {
  int nat = 0, nG = 0;
  const double* tau = nullptr;   // nat x 3, Cartesian Bohr
  const double* born = nullptr;  // nat x 9, (atom, i, j)
  const double* eps = nullptr;   // 9, (i, j)
  const double* G = nullptr;     // nG x 3
  const double* cst = nullptr;   // nat x 9, (atom, i, j)
  double norm = 0., gMax = 14.;
};

inline int ceil_div(long long a, long long b) { return int((a + b - 1) / b); }
inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Mesh point (Gamma-centred) -> integer coordinates (points.cpp:233-240).
__host__ __device__ inline void mesh_coords(int index, int m0, int m1, int& x, int& y, int& z) {
  z = index / (m0 * m1);
  int r = index - z * m0 * m1;
  y = r / m0;
  x = ((r % m0) + m0) % m0;
}

// Folded index of k2 - k1 on the same Gamma-centred mesh (both on the mesh).
__host__ __device__ inline int mesh_diff_index(int i2, int i1, int m0, int m1, int m2) {
  int x1, y1, z1, x2, y2, z2;
  mesh_coords(i1, m0, m1, x1, y1, z1);
  mesh_coords(i2, m0, m1, x2, y2, z2);
  int dx = ((x2 - x1) % m0 + m0) % m0;
  int dy = ((y2 - y1) % m1 + m1) % m1;
  int dz = ((z2 - z1) % m2 + m2) % m2;
  return dz * m0 * m1 + dy * m0 + dx;
}

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet, infinities exact).
__host__ __device__ inline short bf16_bits(float x) {
  union { float f; unsigned u; } v{x};
  unsigned u = v.u;
  if ((u & 0x7f800000u) == 0x7f800000u) return short((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return short(u >> 16);
}
__host__ __device__ inline float bf16_value(short h) {
  union { unsigned u; float f; } v{unsigned(static_cast<unsigned short>(h)) << 16};
  return v.f;
}

}  // namespace phx
