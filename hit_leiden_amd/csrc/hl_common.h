// hl_common.h -- shared types, error handling and device buffers for the
// gfx950 HIT-Leiden engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <vector>

namespace hl {

using u8 = uint8_t;
using u16 = uint16_t;
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
// Same-address device atomics serialise (about 11 ns each on MI355X), so a
// kernel that ends in one atomic per workgroup keeps its grid near this size.
constexpr u32 kReduceBlocks = 256;

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

// Small device-to-host readbacks (counts, maxima, flags) go through a pinned
// staging area: with a pageable destination hipMemcpyAsync waits for its own
// round trip.  d2h_async enqueues the copy on the stream; stream_sync waits for
// the stream and then delivers that stream's staged values to their
// destinations, which must stay alive until then.
void d2h_async(void* dst, const void* src, size_t bytes, hipStream_t s);
// Several small readbacks at once: one kernel writes them all into the pinned
// staging area (one launch in place of one copy each).  Each item is whole
// 4-byte words, 4-byte aligned, at most kD2HGroupBytes; at most kD2HGroupMax items.
struct D2H {
  void* dst;
  const void* src;
  size_t bytes;
};
constexpr size_t kD2HGroupMax = 8, kD2HGroupBytes = 1024;
void d2h_group(std::initializer_list<D2H> items, hipStream_t s);
void stream_sync(hipStream_t s);
// Waits until the values read back on stream s (d2h_async, d2h_group) have
// arrived and delivers them, like stream_sync, but not for work enqueued on s
// after them: the gather kernel stores a token after its values, and the host
// polls the token instead of waiting for the stream (about 5 us less per wait).
// For waits whose only purpose is the values read back; a read-back of a
// runtime copy, or a token that does not come within a second (a failed
// stream), takes stream_sync.
void stream_wait_reads(hipStream_t s);
// Forget the values staged for stream s without delivering them, after
// waiting (errors ignored) for their copies: a call that ended with an
// exception leaves destinations in the frames it unwound.
void drop_staged(hipStream_t s);

// Device allocations made by the calling thread (the engine call in progress):
// how many, their bytes, and the host seconds spent in hipMalloc / hipFree and
// the synchronisation a free needs.  An engine call resets it when it starts
// and reports it in its batch counters.
struct AllocStats {
  u64 count = 0;
  u64 bytes = 0;
  double seconds = 0.0;
  u64 syncs = 0;  // host waits on the engine stream (stream_sync)
};
extern thread_local AllocStats g_alloc;
double host_seconds();
// with HL_ALLOC_TRACE=<MiB> in the environment, every allocation of at least
// that many MiB is printed to stderr with its call site (library offset)
void alloc_trace(size_t bytes, double seconds, const void* site);

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
  __attribute__((noinline)) void reserve(size_t n, hipStream_t s, bool keep = false, double growth = 1.5) {
    if (n <= cap) return;
    const double t0 = host_seconds();
    size_t nc = size_t(double(n) * growth) + 64;
    T* np = nullptr;
    HL_CHECK(hipMalloc(&np, nc * sizeof(T)));
    if (keep && p && cap) HL_CHECK(hipMemcpyAsync(np, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (p) {
      stream_sync(s);
      HL_CHECK(hipFree(p));
    }
    p = np;
    cap = nc;
    const double dt = host_seconds() - t0;
    g_alloc.count += 1;
    g_alloc.bytes += nc * sizeof(T);
    g_alloc.seconds += dt;
    alloc_trace(nc * sizeof(T), dt, __builtin_return_address(0));
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

// ---- optional per-kernel timing with HIP events on the engine stream ----
enum ProfSlot {
  P_ROW_MERGE, P_RESOLVE, P_ROW_DEGREE, P_PROP_SMALL, P_PROP_WAVE, P_PROP_BIG, P_DECOUPLE, P_APPLY_MOVES,
  P_MARK_MOVED, P_LABEL_TOTALS, P_SELECT, P_SORT, P_SCAN, P_CC_SPLIT, P_REFINE, P_AGG_EMIT, P_COMPRESS,
  P_DEF_UPDATE, P_FILL, P_DET_MOVE, P_DET_REFINE, P_TWICE_M, P_SEED, P_MODULARITY, P_REDUCE, P_DEC_MULTI,
  P_PROP_LDS, P_TRY_HALF, P_TRY_WAVE, P_TRY_LDS, P_TRY_HUB, P_TRY_GIANT, P_PROP_TABLE, P_CACHE_UPDATE, P_RELABEL,
  P_MOVE_ROUNDS, P_COUNT
};
inline const char* prof_name(int s) {
  static const char* names[P_COUNT] = {
      "row_merge", "resolve_runs", "row_degree", "propose_small", "propose_wave", "propose_big", "decouple",
      "apply_moves", "mark_moved", "label_totals", "select", "sort", "scan", "cc_split", "refine", "agg_emit",
      "compress", "def_update", "fill", "det_move", "det_refine", "twice_m", "seed", "modularity", "reduce",
      "decouple_multi", "propose_lds", "try_half", "try_wave", "try_lds", "try_hub", "try_giant", "propose_table",
      "cache_update", "relabel", "move_rounds"};
  return (s >= 0 && s < P_COUNT) ? names[s] : "?";
}

struct Profiler {
  bool on = false;
  int only = -1;  // >= 0: record this slot only
  double ms[P_COUNT] = {};
  u64 launches[P_COUNT] = {};
  u64 bytes[P_COUNT] = {};
  struct Pending {
    hipEvent_t a, b;
    int slot;
    u64 bytes;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Pending> pending;
  hipEvent_t get() {
    if (pool.empty()) {
      hipEvent_t e;
      HL_CHECK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  bool records(int sl) const { return on && (only < 0 || only == sl); }
  void reset() {
    for (int i = 0; i < P_COUNT; ++i) ms[i] = 0.0, launches[i] = 0, bytes[i] = 0;
  }
  // waits for the recorded events (a wait for read-back values alone leaves
  // the stream's last kernels running)
  void flush() {
    for (auto& p : pending) {
      float t = 0.f;
      HL_CHECK(hipEventSynchronize(p.b));
      HL_CHECK(hipEventElapsedTime(&t, p.a, p.b));
      ms[p.slot] += t;
      launches[p.slot] += 1;
      bytes[p.slot] += p.bytes;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~Profiler() {
    for (auto& p : pending) pool.push_back(p.a), pool.push_back(p.b);
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
extern thread_local Profiler* g_prof;

struct ProfScope {
  hipStream_t s;
  int slot;
  u64 bytes;
  hipEvent_t a = nullptr;
  ProfScope(int sl, u64 b, hipStream_t st) : s(st), slot(sl), bytes(b) {
    if (g_prof && g_prof->on && (g_prof->only < 0 || g_prof->only == sl)) {
      a = g_prof->get();
      HL_CHECK(hipEventRecord(a, s));
    }
  }
  void set_bytes(u64 b) { bytes = b; }
  ~ProfScope() {
    if (a) {
      hipEvent_t b = g_prof->get();
      if (hipEventRecord(b, s) == hipSuccess) g_prof->pending.push_back({a, b, slot, bytes});
    }
  }
};
#define HL_PROF_CAT2(a, b) a##b
#define HL_PROF_CAT(a, b) HL_PROF_CAT2(a, b)
#define HL_PROF(slot, bytes, stream) ::hl::ProfScope HL_PROF_CAT(prof_scope_, __LINE__)(slot, bytes, stream)

inline unsigned blocks_for(u64 n, unsigned bs = 256, u64 max_blocks = 1u << 20) {
  u64 b = (n + bs - 1) / bs;
  if (b == 0) b = 1;
  if (b > max_blocks) b = max_blocks;
  return unsigned(b);
}

}  // namespace hl
