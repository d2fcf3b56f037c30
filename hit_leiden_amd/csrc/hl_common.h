// hl_common.h -- shared types, error handling and device buffers for the
// gfx950 HIT-Leiden engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace hl {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;

// One edge-delta record (u, v, alpha): GraphInput.edges entries (types.rs:133-148).
struct Rec {
  u32 u, v;
  double w;
};

constexpr u32 kNone = 0xFFFFFFFFu;
constexpr double kEpsApply = 1e-12;    // in_memory.rs:110
constexpr double kEpsSelf = 1e-12;     // hit_leiden.rs:1687
constexpr double kEpsCompress = 1e-9;  // hit_leiden.rs:1711
constexpr double kEpsTFilter = 1e-9;   // hit_leiden.rs:1547,1631
constexpr u32 kSumChunk = 4096;        // canonical chunk of the 2m sum

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const char* what, const char* file, int line)
      : std::runtime_error(std::string(hipGetErrorString(c)) + " at " + what + " (" + file + ":" +
                           std::to_string(line) + ")"),
        code(c) {}
};

struct InputError : std::runtime_error {
  explicit InputError(const std::string& s) : std::runtime_error(s) {}
};

#define HL_CHECK(x)                                                            \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) throw ::hl::HipError(e_, #x, __FILE__, __LINE__);    \
  } while (0)

#define HL_LAUNCH_CHECK() HL_CHECK(hipGetLastError())

// Grow-only device buffer.  reserve(n, keep) reallocates when the capacity is
// short; keep copies the old contents (stream-ordered).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  void reserve(size_t n, hipStream_t s, bool keep = false, double growth = 1.5) {
    if (n <= cap) return;
    size_t nc = size_t(double(n) * growth) + 64;
    T* np = nullptr;
    HL_CHECK(hipMalloc(&np, nc * sizeof(T)));
    if (keep && p && cap) HL_CHECK(hipMemcpyAsync(np, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (p) {
      HL_CHECK(hipStreamSynchronize(s));
      HL_CHECK(hipFree(p));
    }
    p = np;
    cap = nc;
  }
  T* get() const { return p; }
  void swap(DevBuf& o) {
    T* tp = p;
    p = o.p;
    o.p = tp;
    size_t tc = cap;
    cap = o.cap;
    o.cap = tc;
  }
};

inline unsigned blocks_for(u64 n, unsigned bs = 256, u64 max_blocks = 1u << 20) {
  u64 b = (n + bs - 1) / bs;
  if (b == 0) b = 1;
  if (b > max_blocks) b = max_blocks;
  return unsigned(b);
}

}  // namespace hl
