// hl_kernels.hip -- gfx950 kernels of the HIT-Leiden incremental update.
//
// Layout: one CSR per level with rows kept sorted by neighbour id.  Rows that a
// batch touches are rewritten into freshly bump-allocated pool space (the pool
// is compacted when full), so insert/delete never shifts a row in place and
// every row stays canonical for the deterministic mode.
//
// All floating-point gain formulas keep the reference's operation order and are
// compiled with -ffp-contract=off, so a gain here equals the CPU's bit for bit
// for the same inputs.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "hl_kernels.h"
#include "hl_move.h"

namespace hl {
thread_local Profiler* g_prof = nullptr;
thread_local AllocStats g_alloc;

double host_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void alloc_trace(size_t bytes, double seconds, const void* site) {
  static const long min_mib = [] {
    const char* e = std::getenv("HL_ALLOC_TRACE");
    return e ? std::atol(e) : -1L;
  }();
  if (min_mib < 0 || bytes < (size_t(min_mib) << 20)) return;
  Dl_info di{};
  const uintptr_t base = dladdr(site, &di) ? uintptr_t(di.dli_fbase) : 0;
  std::fprintf(stderr, "[hl alloc] %.1f MiB in %.3f ms at +0x%lx\n", double(bytes) / (1 << 20), seconds * 1e3,
               (unsigned long)(uintptr_t(site) - base));
}

namespace {
struct Staging {
  u8* buf = nullptr;
  u8* dbuf = nullptr;  // the area's device address
  size_t cap = 0, used = 0;
  struct Item {
    void* dst;
    size_t off, bytes;
    hipStream_t s;
    u64 seq;         // its gather kernel's call number (0: a runtime copy)
    size_t tok_off;  // where that kernel stores its token, the low 32 bits of seq
  };
  u64 next_seq = 0;
  std::vector<Item> items;
  ~Staging() {
    if (buf) (void)hipHostFree(buf);
  }
};
Staging& staging() {
  thread_local Staging st;
  return st;
}
}  // namespace

__attribute__((noinline)) void stream_sync(hipStream_t s) {
  HL_CHECK(hipStreamSynchronize(s));
  g_alloc.syncs += 1;
  static const bool trace = std::getenv("HL_SYNC_TRACE") != nullptr;
  if (trace) {
    // with HL_SYNC_TRACE set, every wait is printed with its call site (library offset)
    Dl_info di{};
    const void* site = __builtin_return_address(0);
    const uintptr_t base = dladdr(site, &di) ? uintptr_t(di.dli_fbase) : 0;
    std::fprintf(stderr, "[hl sync] +0x%lx\n", (unsigned long)(uintptr_t(site) - base));
  }
  Staging& st = staging();
  size_t keep = 0;
  for (size_t i = 0; i < st.items.size(); ++i) {
    const auto& it = st.items[i];
    if (it.s == s)
      std::memcpy(it.dst, st.buf + it.off, it.bytes);
    else
      st.items[keep++] = it;
  }
  st.items.resize(keep);
  if (keep == 0) st.used = 0;
}

void drop_staged(hipStream_t s) {
  Staging& st = staging();
  bool any = false;
  for (const auto& it : st.items) any = any || it.s == s;
  if (!any) return;
  (void)hipStreamSynchronize(s);  // no copy may land in the area once it is reused
  size_t keep = 0;
  for (size_t i = 0; i < st.items.size(); ++i)
    if (st.items[i].s != s) st.items[keep++] = st.items[i];
  st.items.resize(keep);
  if (keep == 0) st.used = 0;
}

// room for `bytes` in the pinned staging area (stream s about to write it)
static size_t staging_reserve(size_t bytes) {
  Staging& st = staging();
  size_t off = (st.used + 15) & ~size_t(15);
  if (off + bytes > st.cap) {
    if (!st.items.empty()) {
      // deliver what is staged (every stream with pending values), then reuse the area
      std::vector<hipStream_t> ss;
      for (const auto& it : st.items)
        if (std::find(ss.begin(), ss.end(), it.s) == ss.end()) ss.push_back(it.s);
      for (hipStream_t x : ss) stream_sync(x);
    }
    off = 0;
    if (bytes > st.cap) {
      // a gather kernel whose values were taken by their token may not have
      // ended yet: no kernel may run while the area goes
      if (st.buf) HL_CHECK(hipDeviceSynchronize());
      if (st.buf) HL_CHECK(hipHostFree(st.buf));
      st.buf = st.dbuf = nullptr;
      const size_t cap = std::max<size_t>(bytes, 1 << 16);
      // the host polls tokens the gather kernel stores here: coherent memory
      HL_CHECK(hipHostMalloc((void**)&st.buf, cap, hipHostMallocCoherent));
      HL_CHECK(hipHostGetDevicePointer((void**)&st.dbuf, st.buf, 0));
      st.cap = cap;
    }
  }
  return off;
}

void d2h_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes && bytes % 4 == 0 && bytes <= kD2HGroupBytes && reinterpret_cast<uintptr_t>(src) % 4 == 0) {
    // a small readback: the gather kernel, which costs less than a runtime copy
    d2h_group({{dst, src, bytes}}, s);
    return;
  }
  Staging& st = staging();
  const size_t off = staging_reserve(bytes);
  HL_CHECK(hipMemcpyAsync(st.buf + off, src, bytes, hipMemcpyDeviceToHost, s));
  st.items.push_back({dst, off, bytes, s, 0, 0});
  st.used = off + bytes;
}

// the items of a d2h_group, written into the pinned staging area by one kernel
struct D2HGather {
  const u32* src[kD2HGroupMax];
  u32 words[kD2HGroupMax];
  u32 off[kD2HGroupMax];  // in words from the area's start
  u32 n;
};
// one wavefront: its stores of the values, a system-scope fence, then the token
__global__ __launch_bounds__(64) void k_d2h_gather(D2HGather g, u32* area, u32 tok_word, u32 tok) {
  for (u32 i = 0; i < g.n; ++i)
    for (u32 w = threadIdx.x; w < g.words[i]; w += blockDim.x) area[g.off[i] + w] = g.src[i][w];
  __threadfence_system();
  if (threadIdx.x == 0) __hip_atomic_store(area + tok_word, tok, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void d2h_group(std::initializer_list<D2H> items, hipStream_t s) {
  if (items.size() == 0) return;
  if (items.size() > kD2HGroupMax) throw std::logic_error("d2h_group: more than kD2HGroupMax items");
  Staging& st = staging();
  size_t total = 0;
  for (const D2H& it : items) {
    if (it.bytes % 4 || it.bytes > kD2HGroupBytes || reinterpret_cast<uintptr_t>(it.src) % 4)
      throw std::logic_error("d2h_group: items must be 4-byte words, at most kD2HGroupBytes each");
    total += it.bytes;
  }
  const size_t off = staging_reserve(total + 4);  // the values, then the token
  const size_t tok_off = off + total;
  do ++st.next_seq;
  while (u32(st.next_seq) == 0);  // a token is never 0, the value the slot is cleared to
  const u64 seq = st.next_seq;
  const u32 tok = u32(seq);
  // a slot reused from an earlier call holds an older token (or data): clear it
  __atomic_store_n(reinterpret_cast<u32*>(st.buf + tok_off), 0u, __ATOMIC_RELEASE);
  D2HGather g{};
  size_t at = off;
  for (const D2H& it : items) {
    g.src[g.n] = static_cast<const u32*>(it.src);
    g.words[g.n] = u32(it.bytes / 4);
    g.off[g.n] = u32(at / 4);
    ++g.n;
    st.items.push_back({it.dst, at, it.bytes, s, seq, tok_off});
    at += it.bytes;
  }
  hipLaunchKernelGGL(k_d2h_gather, 1, 64, 0, s, g, reinterpret_cast<u32*>(st.dbuf), u32(tok_off / 4), tok);
  HL_LAUNCH_CHECK();
  st.used = off + total + 4;
}

void stream_wait_reads(hipStream_t s) {
  Staging& st = staging();
  // the stream's latest gather (stream order makes its token the last of the
  // stream's values to land); any runtime copy among them: wait for the stream
  u64 seq = 0;
  size_t tok_off = 0;
  bool copies = false, any = false;
  for (const auto& it : st.items) {
    if (it.s != s) continue;
    any = true;
    if (it.seq == 0)
      copies = true;
    else if (it.seq > seq) {
      seq = it.seq;
      tok_off = it.tok_off;
    }
  }
  const u32 tok = u32(seq);
  if (!any || copies) {
    stream_sync(s);
    return;
  }
  const volatile u32* slot = reinterpret_cast<const volatile u32*>(st.buf + tok_off);
  bool arrived = false;
  const double t0 = host_seconds();
  for (u64 spin = 0;; ++spin) {
    if (__atomic_load_n(const_cast<const u32*>(slot), __ATOMIC_ACQUIRE) == tok) {
      arrived = true;
      break;
    }
    if ((spin & 1023) == 1023 && host_seconds() - t0 > 1.0) break;
  }
  if (!arrived) {
    stream_sync(s);  // reports the stream's error, or waits for a slow one
    return;
  }
  g_alloc.syncs += 1;
  size_t keep = 0;
  for (size_t i = 0; i < st.items.size(); ++i) {
    const auto& it = st.items[i];
    if (it.s == s)
      std::memcpy(it.dst, st.buf + it.off, it.bytes);
    else
      st.items[keep++] = it;
  }
  st.items.resize(keep);
  if (keep == 0) st.used = 0;
}
namespace {

constexpr u8 OP_UPD = 1, OP_DEL = 2, OP_INS = 3;
constexpr u32 MERGE_CHUNK = 1024;
constexpr u32 SMALL_MAX = 16;   // thread-per-vertex bucket capacity
constexpr u32 WAVE_HW = 1024;   // LDS hash slots per wavefront
constexpr u32 DET_LCAP = 4096;  // LDS hash slots of the deterministic sweeps
constexpr u32 kDecUndec = 0, kDecAcc = 1, kDecRej = 2;

__device__ __forceinline__ u64 umin64(u64 a, u64 b) { return a < b ? a : b; }
__device__ __forceinline__ u64 gtid() { return u64(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ __forceinline__ u64 gstride() { return u64(gridDim.x) * blockDim.x; }

// hit_leiden.rs:1135-1140 / throughput.rs:335-340 (same operation order)
__device__ __forceinline__ double move_gain(double wc, double wcur, double t, double gamma, double k, double sc,
                                            double scand) {
  return (wc - wcur) / t + gamma * k * (sc - k - scand) / (t * t);
}
// hit_leiden.rs:1150-1154, parallel_frontier.rs:207-210
__device__ __forceinline__ double empty_gain(double wcur, double t, double gamma, double k, double sc) {
  return -wcur / t + gamma * k * (sc - k) / (t * t);
}
// hit_leiden.rs:1546, :1630
__device__ __forceinline__ double remove_gain(double ws, double t, double gamma, double ds, double dp) {
  return -ws / t + gamma * ds * (dp - ds) / (t * t);
}

__device__ __forceinline__ u32 hash32(u32 x) { return x * 2654435761u; }
// T[key] += x in an LDS open-addressing table of cap slots (linear probing;
// the caller keeps the number of distinct keys below cap)
__device__ __forceinline__ void lds_add_keyed(u32* K, double* V, u32 cap, u32 key, double x) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    const u32 old = atomicCAS(&K[h], kNone, key);
    if (old == kNone || old == key) break;
    h = (h + 1) & (cap - 1);
  }
  atomicAdd(&V[h], x);
}

__device__ __forceinline__ u32 lower_bound_u32(const u32* a, u32 n, u32 x) {
  u32 lo = 0, hi = n;
  while (lo < hi) {
    u32 mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// (gain, key) argmax with ties to the smaller key
__device__ __forceinline__ void better(double& g, u32& k, double og, u32 ok) {
  if (og > g || (og == g && ok < k)) {
    g = og;
    k = ok;
  }
}

__device__ __forceinline__ void wave_argmax(double& g, u32& k) {
  for (int o = 32; o > 0; o >>= 1) {
    double og = __shfl_xor(g, o, 64);
    u32 ok = __shfl_xor(k, o, 64);
    better(g, k, og, ok);
  }
}

// block-wide argmax, result valid in every thread (blockDim multiple of 64, <= 1024)
__device__ void block_argmax(double& g, u32& k) {
  __shared__ double sg[16];
  __shared__ u32 sk[16];
  wave_argmax(g, k);
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) {
    sg[wv] = g;
    sk[wv] = k;
  }
  __syncthreads();
  g = sg[0];
  k = sk[0];
  for (u32 i = 1; i < nw; ++i) better(g, k, sg[i], sk[i]);
  __syncthreads();
}

__device__ double block_sum(double x) {
  __shared__ double s[16];
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) s[wv] = x;
  __syncthreads();
  double t = 0.0;
  for (u32 i = 0; i < nw; ++i) t += s[i];
  __syncthreads();
  return t;
}

// ------------------------------------------------------------------ generic
__global__ void k_fill_u8(u8* p, u8 v, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) p[i] = v;
}
__global__ void k_fill_u32(u32* p, u32 v, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) p[i] = v;
}
__global__ void k_fill_u64(u64* p, u64 v, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) p[i] = v;
}
__global__ void k_iota_u32(u32* p, u32 base, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) p[i] = base + u32(i);
}
__global__ void k_copy_u8(u8* d, const u8* s, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) d[i] = s[i];
}
// grid-stride max with one atomic per workgroup; 16-byte loads where the
// array is aligned (device allocations are)
__global__ __launch_bounds__(256) void k_max_u32(const u32* a, u64 n, u32* out) {
  __shared__ u32 sm[4];
  u32 m = 0;
  const bool vec = (reinterpret_cast<uintptr_t>(a) & 15) == 0;
  const u64 nv = vec ? n / 4 : 0;
  const uint4* a4 = reinterpret_cast<const uint4*>(a);
  for (u64 i = gtid(); i < nv; i += gstride()) {
    const uint4 q = a4[i];
    m = max(m, max(max(q.x, q.y), max(q.z, q.w)));
  }
  for (u64 i = 4 * nv + gtid(); i < n; i += gstride()) m = max(m, a[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (u32)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max(max(sm[0], sm[1]), max(sm[2], sm[3]));
    if (m) atomicMax(out, m);
  }
}
// the maxima of up to four arrays of one length in one launch: out[i] takes
// the maximum of a[i] (null arrays are skipped; two may share an output)
struct MaxArrays {
  const u32* a[4];
  u32* out[4];
};
__global__ __launch_bounds__(256) void k_max_multi(MaxArrays g, u64 n) {
  __shared__ u32 sm[4][4];
  u32 m[4] = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const u32* a = g.a[k];
    if (!a) continue;
    const bool vec = (reinterpret_cast<uintptr_t>(a) & 15) == 0;
    const u64 nv = vec ? n / 4 : 0;
    const uint4* a4 = reinterpret_cast<const uint4*>(a);
    for (u64 i = gtid(); i < nv; i += gstride()) {
      const uint4 q = a4[i];
      m[k] = max(m[k], max(max(q.x, q.y), max(q.z, q.w)));
    }
    for (u64 i = 4 * nv + gtid(); i < n; i += gstride()) m[k] = max(m[k], a[i]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    for (int o = 32; o > 0; o >>= 1) m[k] = max(m[k], (u32)__shfl_xor((int)m[k], o, 64));
    if ((threadIdx.x & 63) == 0) sm[k][threadIdx.x >> 6] = m[k];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = int(threadIdx.x);
    const u32 x = max(max(sm[k][0], sm[k][1]), max(sm[k][2], sm[k][3]));
    if (g.a[k] && x) atomicMax(g.out[k], x);
  }
}
// Ordered stream compaction of a byte-flag array into the ascending list of
// flagged indices, in two launches and without a look-back: each workgroup owns
// a chunk of kSelChunk flags (kSelPer consecutive flags per thread).
constexpr u32 kSelPer = 16, kSelChunk = 256 * kSelPer;
__device__ __forceinline__ u32 sel_load(const u8* flags, u64 n, u64 i0, u8 (&f)[kSelPer]) {
  u32 c = 0;
  if (i0 + kSelPer <= n && ((reinterpret_cast<uintptr_t>(flags + i0) & 15) == 0)) {
    const uint4 q = *reinterpret_cast<const uint4*>(flags + i0);
    const u32 wds[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (u32 j = 0; j < kSelPer; ++j) f[j] = u8(wds[j >> 2] >> (8 * (j & 3)));
  } else {
#pragma unroll
    for (u32 j = 0; j < kSelPer; ++j) f[j] = (i0 + j < n) ? flags[i0 + j] : u8(0);
  }
#pragma unroll
  for (u32 j = 0; j < kSelPer; ++j) c += f[j] != 0;
  return c;
}
// block-wide sum of x (256 threads); every thread gets the total
__device__ __forceinline__ u32 sel_block_sum(u32 x, u32* sm) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = x;
  __syncthreads();
  return sm[0] + sm[1] + sm[2] + sm[3];
}
__global__ __launch_bounds__(256) void k_sel_count(const u8* flags, u64 n, u32* bcnt) {
  __shared__ u32 sm[4];
  u8 f[kSelPer];
  const u64 i0 = u64(blockIdx.x) * kSelChunk + u64(threadIdx.x) * kSelPer;
  const u32 t = sel_block_sum(sel_load(flags, n, i0, f), sm);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = t;
}
__global__ __launch_bounds__(256) void k_sel_write(const u8* flags, u64 n, const u32* bcnt, u32* out,
                                                   unsigned long long* total) {
  __shared__ u32 wsum[4];
  // this chunk's offset: the counts of the chunks before it
  u64 pre = 0;
  for (u32 b = threadIdx.x; b < blockIdx.x; b += 256) pre += bcnt[b];
  for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o, 64);
  __shared__ u64 psm[4];
  if ((threadIdx.x & 63) == 0) psm[threadIdx.x >> 6] = pre;
  __syncthreads();
  pre = psm[0] + psm[1] + psm[2] + psm[3];
  u8 f[kSelPer];
  const u64 i0 = u64(blockIdx.x) * kSelChunk + u64(threadIdx.x) * kSelPer;
  const u32 c = sel_load(flags, n, i0, f);
  // exclusive prefix of c over the workgroup: wavefront scan, then wave sums
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u32 inc = c;
  for (int o = 1; o < 64; o <<= 1) {
    const u32 y = __shfl_up(inc, o, 64);
    if (lane >= u32(o)) inc += y;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  u32 wbase = 0;
  for (u32 i = 0; i < wv; ++i) wbase += wsum[i];
  u64 pos = pre + wbase + inc - c;
#pragma unroll
  for (u32 j = 0; j < kSelPer; ++j)
    if (f[j]) out[pos++] = u32(i0 + j);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) *total = pos;
}
// counts of nonzero bytes in up to four arrays, one atomic per workgroup and
// array; four flags per 4-byte load where the arrays are aligned
__device__ __forceinline__ u32 nz_bytes(u32 x) {
  return u32((x & 0xFFu) != 0) + u32((x & 0xFF00u) != 0) + u32((x & 0xFF0000u) != 0) + u32((x >> 24) != 0);
}
__global__ __launch_bounds__(256) void k_count4(const u8* a0, const u8* a1, const u8* a2, const u8* a3, u64 n,
                                                 unsigned long long* out) {
  __shared__ unsigned long long sm[4][4];
  const u8* arr[4] = {a0, a1, a2, a3};
  bool vec = true;
  for (int j = 0; j < 4; ++j) vec = vec && (reinterpret_cast<uintptr_t>(arr[j]) & 3) == 0;
  const u64 nv = vec ? n / 4 : 0;
  unsigned long long c[4] = {0, 0, 0, 0};
  for (u64 i = gtid(); i < nv; i += gstride())
    for (int j = 0; j < 4; ++j)
      if (arr[j]) c[j] += nz_bytes(reinterpret_cast<const u32*>(arr[j])[i]);
  for (u64 i = 4 * nv + gtid(); i < n; i += gstride())
    for (int j = 0; j < 4; ++j)
      if (arr[j]) c[j] += (arr[j][i] != 0);
  for (int j = 0; j < 4; ++j) {
    for (int o = 32; o > 0; o >>= 1) c[j] += __shfl_xor(c[j], o, 64);
    if ((threadIdx.x & 63) == 0) sm[j][threadIdx.x >> 6] = c[j];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = sm[threadIdx.x][0] + sm[threadIdx.x][1] + sm[threadIdx.x][2] + sm[threadIdx.x][3];
    if (t) atomicAdd(&out[threadIdx.x], t);
  }
}
__global__ void k_fill4_u8(u8* a0, u8* a1, u8* a2, u8* a3, u8 v, u64 n) {
  for (u64 i = gtid(); i < n; i += gstride()) {
    if (a0) a0[i] = v;
    if (a1) a1[i] = v;
    if (a2) a2[i] = v;
    if (a3) a3[i] = v;
  }
}
__global__ void k_scan_tail(const u64* in, u64* out, u64 n) {
  out[n] = n ? out[n - 1] + in[n - 1] : 0;
}

// ------------------------------------------------------------------ apply delta
// also clears the resolve kernel's update count (one launch less)
__global__ void k_canon_keys(const Rec* d, u64 k, u64* keys, u32* idx, u32* ucount) {
  if (gtid() == 0) *ucount = 0;
  for (u64 i = gtid(); i < k; i += gstride()) {
    Rec r = d[i];
    u32 a = min(r.u, r.v), b = max(r.u, r.v);
    keys[i] = (u64(a) << 32) | b;
    idx[i] = u32(i);
  }
}
__global__ void k_heads_u64(const u64* keys, u64 k, int shift, u8* heads) {
  for (u64 i = gtid(); i < k; i += gstride()) heads[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift));
}

// One thread per canonical key: existing weight (or 0) + alpha_1 + ... in delta
// order (apply_delta_edges, in_memory.rs:133-138); the key survives iff > 1e-12
// (in_memory.rs:110).  Emits one directed update per row end.
__global__ void k_resolve_runs(const unsigned long long* nruns_dev, const u32* runs, u64 k, const u64* keys,
                               const u32* idx, const Rec* d, LevelCsrMut g, u64* ukey, u32* uidx, u8* uop,
                               double* uw, u32* ucount) {
  // nruns from the device (the run selection's count), so the host need not
  // read it back first; outputs appended with one atomic per wavefront
  const u64 nruns = *nruns_dev;
  const u32 lane = threadIdx.x & 63;
  for (u64 base = gtid() - lane; base < nruns; base += gstride()) {  // uniform per wavefront
    const u64 r = base + lane;
    u32 n_out = 0, a = 0, b = 0;
    u8 op = 0;
    double acc = 0.0;
    if (r < nruns) {
      const u64 s = runs[r], e = (r + 1 < nruns) ? runs[r + 1] : k;
      const u64 key = keys[s];
      a = u32(key >> 32);
      b = u32(key & 0xffffffffu);
      const u64 off = g.row_off[a];
      const u32 len = g.row_len[a];
      const u32 pos = lower_bound_u32(g.col + off, len, b);
      const bool found = pos < len && g.col[off + pos] == b;
      acc = found ? g.w[off + pos] : 0.0;
      for (u64 t = s; t < e; ++t) acc += d[idx[t]].w;
      if (acc > kEpsApply) {
        op = found ? OP_UPD : OP_INS;
        n_out = (a == b) ? 1u : 2u;
      } else if (found) {
        op = OP_DEL;
        n_out = (a == b) ? 1u : 2u;
      }
    }
    u32 incl = n_out;
    for (int o = 1; o < 64; o <<= 1) {
      const u32 y = __shfl_up(incl, o, 64);
      if (lane >= u32(o)) incl += y;
    }
    const u32 tot = __shfl(incl, 63, 64);
    u32 wbase = 0;
    if (lane == 0 && tot) wbase = atomicAdd(ucount, tot);
    wbase = __shfl(wbase, 0, 64);
    if (n_out) {
      const u32 at = wbase + incl - n_out;
      ukey[at] = (u64(a) << 32) | b;
      uop[at] = op;
      uw[at] = acc;
      uidx[at] = at;
      if (n_out == 2) {
        ukey[at + 1] = (u64(b) << 32) | a;
        uop[at + 1] = op;
        uw[at + 1] = acc;
        uidx[at + 1] = at + 1;
      }
    }
  }
}

// Per sorted update: (is insert) << 32 | (is delete), for one scan that gives
// every update its inserts and deletes before it (a row's updates are
// contiguous, so the row's share is a difference of two scan entries).
// also clears the row plan's three counters (one launch less)
__global__ void k_upd_code(u64 nupd, const u32* perm, const u8* uop, u64* code, u64* plan_cnt) {
  if (gtid() < 3) plan_cnt[gtid()] = 0;
  for (u64 t = gtid(); t < nupd; t += gstride()) {
    const u8 op = uop[perm[t]];
    code[t] = (u64(op == OP_INS) << 32) | u64(op == OP_DEL);
  }
}
__global__ void k_upd_pre(u64 nupd, u64 nseg, const u32* segs, const u64* scan, u32* ins_pre, u32* del_pre) {
  for (u64 t = gtid(); t < nupd; t += gstride()) {
    u64 lo = 0, hi = nseg;  // the last segment starting at or before t
    while (hi - lo > 1) {
      const u64 mid = (lo + hi) >> 1;
      if (segs[mid] <= t) lo = mid;
      else hi = mid;
    }
    const u64 d = scan[t] - scan[segs[lo]];
    ins_pre[t] = u32(d >> 32);
    del_pre[t] = u32(d);
  }
}
// Pool room of a row that outgrows its capacity: an eighth more than its new
// length, and at least two entries, so that the next batches' inserts into it
// are merged in place.
__device__ __forceinline__ u32 row_slack(u32 nl) { return nl ? nl / 8 + 2 : 0; }

// Per updated row: its counts, whether its new length fits its capacity (then
// it is rewritten in place from a staged copy of its old entries: stlen = old
// length, newcap = 0) or it moves to new pool space of newcap entries.
__global__ void k_row_plan(u64 nseg, const u32* segs, u64 nupd, const u64* ukey_s, const u64* scan,
                           const u32* row_len, const u32* row_cap, u32* seg_row, u32* seg_len, u32* new_len,
                           u32* seg_ins, u32* seg_del, u64* newcap, u64* stlen, u64* chunks, u64* old_sum) {
  // old_sum[0] += old row lengths, old_sum[1] = max new length, old_sum[2] +=
  // new row lengths: per-thread partials, one atomic each per workgroup
  unsigned long long osum = 0, omax = 0, nsum = 0;
  for (u64 j = gtid(); j < nseg; j += gstride()) {
    u64 s = segs[j], e = (j + 1 < nseg) ? segs[j + 1] : nupd;
    u32 r = u32(ukey_s[s] >> 32);
    const u64 d = scan[e] - scan[s];
    const u32 ins = u32(d >> 32), del = u32(d);
    u32 old = row_len[r];
    u32 nl = old + ins - del;
    const bool fits = nl <= row_cap[r];
    seg_row[j] = r;
    seg_len[j] = u32(e - s);
    new_len[j] = nl;
    seg_ins[j] = ins;
    seg_del[j] = del;
    newcap[j] = fits ? 0 : u64(nl) + row_slack(nl);
    stlen[j] = fits ? old : 0;
    u64 items = u64(old) + (e - s);
    chunks[j] = (items + MERGE_CHUNK - 1) / MERGE_CHUNK;
    osum += old;
    nsum += nl;
    omax = omax > nl ? omax : nl;
  }
  __shared__ unsigned long long ps[4], pm[4], pn[4];
  for (int o = 32; o > 0; o >>= 1) {
    osum += __shfl_xor(osum, o, 64);
    nsum += __shfl_xor(nsum, o, 64);
    const unsigned long long y = __shfl_xor(omax, o, 64);
    omax = omax > y ? omax : y;
  }
  if ((threadIdx.x & 63) == 0) {
    ps[threadIdx.x >> 6] = osum;
    pm[threadIdx.x >> 6] = omax;
    pn[threadIdx.x >> 6] = nsum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0, m = 0, q = 0;
    for (u32 w = 0; w < blockDim.x / 64; ++w) {
      t += ps[w];
      q += pn[w];
      m = m > pm[w] ? m : pm[w];
    }
    if (t) atomicAdd((unsigned long long*)old_sum, t);
    if (m) atomicMax((unsigned long long*)(old_sum + 1), m);
    if (q) atomicAdd((unsigned long long*)(old_sum + 2), q);
  }
}

// chunk_seg[b] = the segment j of merge block b (chunkbase[j] <= b <
// chunkbase[j+1]), one thread per segment: the stage and merge kernels then
// find their segment with one load instead of a binary search over the
// chunk bases (a chain of dependent loads in every block).  When every row
// is one chunk (the common case) there is no table: block b is segment b.
__global__ void k_chunk_seg(u64 nseg, const u64* chunkbase, u32* chunk_seg) {
  for (u64 j = gtid(); j < nseg; j += gstride())
    for (u64 b = chunkbase[j]; b < chunkbase[j + 1]; ++b) chunk_seg[b] = u32(j);
}

// The old entries of the rows updated in place, copied aside (the same
// block-to-chunk mapping as the merge), so that the merge never reads what it
// overwrites.
__global__ __launch_bounds__(256) void k_row_stage(const u32* chunk_seg, const u64* chunkbase, const u32* seg_row,
                                                    const u64* stlen, const u64* stoff, LevelCsrMut g, u32* st_col,
                                                    double* st_w) {
  const u64 b = blockIdx.x;
  const u64 j = chunk_seg ? chunk_seg[b] : b;
  const u64 olen = stlen[j];
  if (olen == 0) return;
  const u64 c = b - chunkbase[j];
  const u64 oo = g.row_off[seg_row[j]], so = stoff[j];
  const u64 end = umin64(olen, (c + 1) * MERGE_CHUNK);
  for (u64 t = c * MERGE_CHUNK + threadIdx.x; t < end; t += blockDim.x) {
    st_col[so + t] = g.col[oo + t];
    st_w[so + t] = g.w[oo + t];
  }
}

// Merge an old sorted row with its sorted updates into its place: its own
// position when it fits its capacity (the old entries read from their staged
// copy), else its new pool space.  Existing entry t goes to t - #deleted-before
// + #inserted-before; an insert goes to lower_bound(old row) - #deleted-before
// + #inserted-before.
__global__ __launch_bounds__(256) void k_row_merge(const u32* chunk_seg, const u64* chunkbase, const u32* segs,
                                                    const u32* seg_row, const u32* seg_len, const u32* seg_ins,
                                                    const u32* seg_del, const u64* ukey_s, const u32* perm,
                                                    const u8* uop, const double* uw, const u32* ins_pre,
                                                    const u32* del_pre, const u64* newoff, const u64* newcap,
                                                    const u64* stoff, const u32* st_col, const double* st_w,
                                                    u64 pool_base, LevelCsrMut g) {
  const u64 j = chunk_seg ? chunk_seg[blockIdx.x] : blockIdx.x;
  const u64 c = blockIdx.x - chunkbase[j];
  const u32 r = seg_row[j];
  const u64 oo = g.row_off[r];
  const u32 olen = g.row_len[r];
  const u64 s = segs[j];
  const u32 ul = seg_len[j];
  const bool inplace = newcap[j] == 0;
  const u32* ocol = inplace ? st_col + stoff[j] : g.col + oo;
  const double* ow = inplace ? st_w + stoff[j] : g.w + oo;
  const u64 dst = inplace ? oo : pool_base + newoff[j];
  const u64 items = u64(olen) + ul;
  const u64 end = umin64(items, (c + 1) * MERGE_CHUNK);
  for (u64 t = c * MERGE_CHUNK + threadIdx.x; t < end; t += blockDim.x) {
    if (t < olen) {
      u32 col = ocol[t];
      // first update with col >= this col
      u32 lo2 = 0, hi2 = ul;
      while (lo2 < hi2) {
        u32 mid = (lo2 + hi2) >> 1;
        if (u32(ukey_s[s + mid] & 0xffffffffu) < col)
          lo2 = mid + 1;
        else
          hi2 = mid;
      }
      u32 q = lo2;
      u32 dp = q < ul ? del_pre[s + q] : seg_del[j];
      u32 ip = q < ul ? ins_pre[s + q] : seg_ins[j];
      bool match = q < ul && u32(ukey_s[s + q] & 0xffffffffu) == col;
      double w = ow[t];
      if (match) {
        u32 pi = perm[s + q];
        u8 op = uop[pi];
        if (op == OP_DEL) continue;
        w = uw[pi];
      }
      u64 pos = t - dp + ip;
      g.col[dst + pos] = col;
      g.w[dst + pos] = w;
    } else {
      u32 q = u32(t - olen);
      u32 pi = perm[s + q];
      if (uop[pi] != OP_INS) continue;
      u32 col = u32(ukey_s[s + q] & 0xffffffffu);
      u32 lb = lower_bound_u32(ocol, olen, col);
      u64 pos = u64(lb) - del_pre[s + q] + ins_pre[s + q];
      g.col[dst + pos] = col;
      g.w[dst + pos] = uw[pi];
    }
  }
}

__global__ void k_row_commit(u64 nseg, const u32* seg_row, const u32* new_len, const u64* newoff, const u64* newcap,
                             u64 pool_base, LevelCsrMut g) {
  for (u64 j = gtid(); j < nseg; j += gstride()) {
    u32 r = seg_row[j];
    g.row_len[r] = new_len[j];
    if (newcap[j]) {
      g.row_off[r] = pool_base + newoff[j];
      g.row_cap[r] = u32(newcap[j]);
    }
  }
}

// k_v = sequential sum over the sorted row, self-loop counted twice (the order
// of the canonical CPU oracle; equal to any order for integer weights)
// When the maintained community / subcommunity totals are live, the change of
// k_v is added to the totals of v's community and subcommunity.
__global__ void k_row_degree(u64 nseg, const u32* seg_row, LevelCsrMut g, TotalsRef tr) {
  for (u64 j = gtid(); j < nseg; j += gstride()) {
    u32 r = seg_row[j];
    u64 off = g.row_off[r];
    u32 len = g.row_len[r];
    double acc = 0.0;
    for (u32 i = 0; i < len; ++i) {
      double w = g.w[off + i];
      acc += w;
      if (g.col[off + i] == r) acc += w;
    }
    const double old = g.k[r];
    g.k[r] = acc;
    if (acc != old) {
      if (tr.tot) atomicAdd(&tr.tot[tr.f[r]], acc - old);
      if (tr.dS) atomicAdd(&tr.dS[tr.s[r]], acc - old);
    }
  }
}

// the new degree of row r (acc), and the totals' share of its change
__device__ __forceinline__ void set_degree(const LevelCsrMut& g, const TotalsRef& tr, u32 r, double acc) {
  const double old = g.k[r];
  g.k[r] = acc;
  if (acc != old) {
    if (tr.tot) atomicAdd(&tr.tot[tr.f[r]], acc - old);
    if (tr.dS) atomicAdd(&tr.dS[tr.s[r]], acc - old);
    if (tr.tm) atomicAdd(&tr.tm[tr.gid[r]], acc - old);
  }
}
// Throughput mode: one wavefront per updated row, a fixed-shape tree sum
// (the deterministic mode keeps the sequential row fold above).  Rows above
// kIntraHub entries go to the hub list when there is one: their sums come from
// pieces over the grid (k_row_degree_hubs), not from the one wavefront.
__global__ __launch_bounds__(256) void k_row_degree_wave(u64 nseg, const u32* seg_row, LevelCsrMut g, TotalsRef tr,
                                                         u32* hubs, u32* nhubs, u32 hub_cap) {
  const u32 lane = threadIdx.x & 63;
  const u64 wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (u64 j = wave; j < nseg; j += nwaves) {
    const u32 r = seg_row[j];
    const u64 off = g.row_off[r];
    const u32 len = g.row_len[r];
    if (hubs && len > kIntraHub) {
      u32 slot = 0;
      if (lane == 0) slot = atomicAdd(nhubs, 1u);
      slot = __shfl(slot, 0, 64);
      if (slot < hub_cap) {
        if (lane == 0) hubs[slot] = r;
        continue;
      }
    }
    double acc = 0.0;
    for (u32 i = lane; i < len; i += 64) {
      const double w = g.w[off + i];
      acc += w;
      if (g.col[off + i] == r) acc += w;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) set_degree(g, tr, r, acc);
  }
}
// the listed rows in pieces of kIntraPiece entries, one workgroup per piece:
// each piece's sum into the row's accumulator
__global__ __launch_bounds__(256) void k_row_degree_hubs(LevelCsrMut g, const u32* hubs, const u32* nhubs,
                                                         u32 hub_cap, u32 pieces, double* acc) {
  const u32 nh = min(*nhubs, hub_cap);
  for (u64 q = blockIdx.x; q < u64(nh) * pieces; q += gridDim.x) {
    const u32 r = hubs[q / pieces];
    const u64 b = (q % pieces) * u64(kIntraPiece);
    const u32 len = g.row_len[r];
    if (b >= len) continue;
    const u64 e = min<u64>(len, b + kIntraPiece), off = g.row_off[r];
    double local = 0.0;
    for (u64 i = b + threadIdx.x; i < e; i += 256) {
      const double w = g.w[off + i];
      local += w;
      if (g.col[off + i] == r) local += w;
    }
    const double t = block_sum(local);
    if (threadIdx.x == 0) atomicAdd(&acc[q / pieces], t);
  }
}
__global__ void k_row_degree_fin(LevelCsrMut g, TotalsRef tr, const u32* hubs, const u32* nhubs, u32 hub_cap,
                                 const double* acc) {
  const u32 nh = min(*nhubs, hub_cap);
  for (u64 i = gtid(); i < nh; i += gstride()) set_degree(g, tr, hubs[i], acc[i]);
}

__global__ void k_copy_rows(LevelCsrMut g, const u64* newoff, u32* col2, double* w2) {
  const u32 lane = threadIdx.x & 63;
  const u64 wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (u64 r = wave; r < g.n; r += nwaves) {
    u64 o = g.row_off[r], d = newoff[r];
    u32 len = g.row_len[r];
    for (u32 i = lane; i < len; i += 64) {
      col2[d + i] = g.col[o + i];
      w2[d + i] = g.w[o + i];
    }
  }
}
__global__ void k_set_offsets(LevelCsrMut g, const u64* newoff) {
  for (u64 r = gtid(); r < g.n; r += gstride()) g.row_off[r] = newoff[r];
}
__global__ void k_len_u64(const u32* len, u32 n, u64* out) {
  for (u64 i = gtid(); i < n; i += gstride()) out[i] = len[i];
}

// canonical 2m: pairwise (halving) tree over k zero-padded to a power of two.
// Each workgroup reduces one aligned 2048-element block with the same tree
// (element i absorbs i+s for s = 1, 2, 4, ...); block results feed the next pass.
constexpr u32 kTreeBlock = 2048;
__global__ __launch_bounds__(256) void k_tree_pass(const double* in, u64 n, double* out) {
  __shared__ double t[kTreeBlock];
  const u64 base = u64(blockIdx.x) * kTreeBlock;
  for (u32 i = threadIdx.x; i < kTreeBlock; i += 256) t[i] = (base + i < n) ? in[base + i] : 0.0;
  __syncthreads();
  for (u32 s = 1; s < kTreeBlock; s <<= 1) {
    for (u32 i = threadIdx.x * 2 * s; i < kTreeBlock; i += 256 * 2 * s) t[i] += t[i + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = t[0];
}

__global__ void k_grow_init(u32 n_old, u32 n_new, u32 bf, u32 bg, u32 bs, u32* f, u32* g, u32* s_pre, u32* s_cur,
                            u64* row_off, u32* row_len, u32* row_cap, double* k) {
  u64 cnt = n_new - n_old;
  for (u64 i = gtid(); i < cnt; i += gstride()) {
    u32 v = n_old + u32(i);
    f[v] = bf + u32(i);
    g[v] = bg + u32(i);
    s_pre[v] = bs + u32(i);
    s_cur[v] = bs + u32(i);
    row_off[v] = 0;
    row_len[v] = 0;
    row_cap[v] = 0;
    k[v] = 0.0;
  }
}

// ------------------------------------------------------------------ seeding
// hit_leiden.rs:1037-1052
__global__ void k_seed(const Rec* d, u64 k, const u32* f, const u32* s, u8* A, u8* K) {
  for (u64 i = gtid(); i < k; i += gstride()) {
    Rec r = d[i];
    bool same_f = f[r.u] == f[r.v];
    if ((r.w > 0.0 && !same_f) || (r.w < 0.0 && same_f)) {
      A[r.u] = 1;
      A[r.v] = 1;
    }
    if (s[r.u] == s[r.v]) {
      K[r.u] = 1;
      K[r.v] = 1;
    }
  }
}

// tot[lab[v]] += k[v], combined per workgroup pass in an LDS hash table (one
// device atomic per distinct label of the pass): the members of one big
// community would otherwise serialise on one address
constexpr u32 kTotTable = 2048, kTotPer = 4;  // kTotPer * 256 insertions never fill the table
__global__ __launch_bounds__(256) void k_label_totals_atomic(const u32* lab, const double* k, u32 n, double* tot) {
  __shared__ u32 tk[kTotTable];
  __shared__ double tv[kTotTable];
  for (u64 base = u64(blockIdx.x) * 256 * kTotPer; base < n; base += u64(gridDim.x) * 256 * kTotPer) {
    for (u32 i = threadIdx.x; i < kTotTable; i += 256) {
      tk[i] = kNone;
      tv[i] = 0.0;
    }
    __syncthreads();
    for (u32 r = 0; r < kTotPer; ++r) {
      const u64 v = base + u64(r) * 256 + threadIdx.x;
      if (v >= n) continue;
      const double x = k[v];
      if (x == 0.0) continue;
      const u32 key = lab[v];
      u32 h = hash32(key) & (kTotTable - 1);
      while (true) {
        const u32 old = atomicCAS(&tk[h], kNone, key);
        if (old == kNone || old == key) break;
        h = (h + 1) & (kTotTable - 1);
      }
      atomicAdd(&tv[h], x);
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < kTotTable; i += 256)
      if (tk[i] != kNone) atomicAdd(&tot[tk[i]], tv[i]);
    __syncthreads();
  }
}
// per label segment (vertices ascending): sequential sum
__global__ void k_segment_sums(u64 nseg, const u32* starts, u64 n, const u32* keys_s, const u32* vals_s,
                               const double* k, double* tot) {
  for (u64 j = gtid(); j < nseg; j += gstride()) {
    u64 s = starts[j], e = (j + 1 < nseg) ? starts[j + 1] : n;
    double acc = 0.0;
    for (u64 t = s; t < e; ++t) acc += k[vals_s[t]];
    tot[keys_s[s]] = acc;
  }
}
__global__ void k_heads_u32(const u32* keys, u64 n, u8* heads) {
  for (u64 i = gtid(); i < n; i += gstride()) heads[i] = (i == 0 || keys[i] != keys[i - 1]);
}

// ------------------------------------------------------------------ deterministic sweep
struct DetMoveArgs {
  Csr g;
  u32* f;
  const u32* s;
  double* tot;
  u64 tot_cap;
  u64* tot_len;
  u32 *bm0, *bm1, *bm2;
  u32 nw2;
  u8* B;
  u8* K;
  double gamma;
  u32* gk;
  double* gv;
  u32* gused;
  u32 gcap;
  u32* status;
  u64* pops;
};

__device__ __forceinline__ void bm_set(u32* bm0, u32* bm1, u32* bm2, u32 v) {
  u32 i0 = v >> 5;
  u32 old = atomicOr(&bm0[i0], 1u << (v & 31));
  if (old == 0) {
    u32 i1 = i0 >> 5;
    u32 o1 = atomicOr(&bm1[i1], 1u << (i0 & 31));
    if (o1 == 0) atomicOr(&bm2[i1 >> 5], 1u << (i1 & 31));
  }
}

__global__ void k_bm_level0(const u8* A, u32 n, u32* bm0, u64 nw0) {
  for (u64 w = gtid(); w < nw0; w += gstride()) {
    u32 x = 0;
    for (u32 b = 0; b < 32; ++b) {
      u64 v = w * 32 + b;
      if (v < n && A[v]) x |= 1u << b;
    }
    bm0[w] = x;
  }
}
__global__ void k_bm_summary(const u32* lo, u64 nlo, u32* hi, u64 nhi) {
  for (u64 w = gtid(); w < nhi; w += gstride()) {
    u32 x = 0;
    for (u32 b = 0; b < 32; ++b) {
      u64 i = w * 32 + b;
      if (i < nlo && lo[i]) x |= 1u << b;
    }
    hi[w] = x;
  }
}

// Open-addressing lookup (linear probing); 0.0 for an absent key.
__device__ __forceinline__ double table_lookup(const u32* K, const double* V, u32 cap, u32 key) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    u32 x = K[h];
    if (x == key) return V[h];
    if (x == kNone) return 0.0;
    h = (h + 1) & (cap - 1);
  }
}
// Sequential insert used by the deterministic sweeps (one thread): keeps the
// per-key sums in row order.
__device__ __forceinline__ u32 seq_insert(u32* K, u32 cap, u32 key, u32* used, u32& nused) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    u32 x = K[h];
    if (x == key) return h;
    if (x == kNone) {
      K[h] = key;
      used[nused++] = h;
      return h;
    }
    h = (h + 1) & (cap - 1);
  }
}

// hit_leiden.rs:1108-1188 in one workgroup: pop the lowest active vertex,
// accumulate w(v,C) in row order (self-loop entries included), argmax over
// communities in ascending id, empty community id = current array length.
__global__ __launch_bounds__(256) void k_det_move(DetMoveArgs a) {
  __shared__ u32 lk[DET_LCAP];
  __shared__ double lv[DET_LCAP];
  __shared__ u32 lused[DET_LCAP];
  __shared__ u32 cl[256];
  __shared__ double cw[256];
  __shared__ u32 s_v, s_done, s_nused;
  __shared__ double s_wcur;
  __shared__ u32 s_best;
  __shared__ double s_bg;
  const u32 tid = threadIdx.x;
  for (u32 q = tid; q < DET_LCAP; q += 256) {
    lk[q] = kNone;
    lv[q] = 0.0;
  }
  if (tid == 0) {
    s_best = kNone;
    s_bg = 0.0;
  }
  __syncthreads();
  const Csr& g = a.g;
  const double tt = *g.twice_m;
  u64 pops = 0;
  while (true) {
    if (tid < 64) {
      u32 found = kNone;
      for (u32 base = 0; base < a.nw2; base += 64) {
        u32 w = (base + tid < a.nw2) ? a.bm2[base + tid] : 0u;
        unsigned long long m = __ballot(w != 0);
        if (m) {
          u32 l = u32(__ffsll((long long)m) - 1);
          u32 word = __shfl(w, l, 64);
          found = (base + l) * 32 + u32(__ffs(word) - 1);
          break;
        }
      }
      if (tid == 0) {
        if (found == kNone) {
          s_done = 1;
        } else {
          u32 i1 = found;
          u32 w1 = a.bm1[i1];
          u32 i0 = i1 * 32 + u32(__ffs(w1) - 1);
          u32 w0 = a.bm0[i0];
          u32 v = i0 * 32 + u32(__ffs(w0) - 1);
          w0 &= ~(1u << (v & 31));
          a.bm0[i0] = w0;
          if (w0 == 0) {
            w1 &= ~(1u << (i0 & 31));
            a.bm1[i1] = w1;
            if (w1 == 0) a.bm2[i1 >> 5] &= ~(1u << (i1 & 31));
          }
          s_v = v;
          s_done = 0;
          s_nused = 0;
        }
      }
    }
    __syncthreads();
    if (s_done) break;
    ++pops;
    const u32 v = s_v;
    const u32 cur = a.f[v];
    const double kv = g.k[v];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const bool in_lds = deg <= DET_LCAP / 2;
    u32* HK = in_lds ? lk : a.gk;
    double* HV = in_lds ? lv : a.gv;
    u32* used = in_lds ? lused : a.gused;
    const u32 cap = in_lds ? DET_LCAP : a.gcap;
    for (u32 base = 0; base < deg; base += 256) {
      u32 j = base + tid;
      if (j < deg) {
        u32 u = g.col[off + j];
        // self-loops skipped, as parallel_frontier.rs:164-166 does
        cl[tid] = (u == v) ? kNone : a.f[u];
        cw[tid] = g.w[off + j];
      }
      __syncthreads();
      if (tid == 0) {
        u32 m = min(256u, deg - base);
        u32 nu = s_nused;
        for (u32 q = 0; q < m; ++q) {
          if (cl[q] == kNone) continue;
          u32 h = seq_insert(HK, cap, cl[q], used, nu);
          HV[h] += cw[q];
        }
        s_nused = nu;
      }
      __syncthreads();
    }
    if (tid == 0) s_wcur = table_lookup(HK, HV, cap, cur);
    __syncthreads();
    const double wcur = s_wcur;
    const double sc = a.tot[cur];
    const u32 nused = s_nused;
    double bg = 0.0;
    u32 best = kNone;
    for (u32 q = tid; q < nused; q += 256) {
      u32 h = used[q];
      u32 key = HK[h];
      if (key == cur) continue;
      double gn = move_gain(HV[h], wcur, tt, a.gamma, kv, sc, a.tot[key]);
      if (gn > 0.0) better(bg, best, gn, key);
    }
    block_argmax(bg, best);
    if (tid == 0) {
      u64 next_empty = *a.tot_len;
      double eg = empty_gain(wcur, tt, a.gamma, kv, sc);
      u64 b64 = best;
      if (eg > bg) {
        bg = eg;
        b64 = next_empty;
      }
      s_bg = bg;
      if (bg > 0.0) {
        if (b64 >= a.tot_cap) {
          *a.status = 1;  // grow and resume
          bm_set(a.bm0, a.bm1, a.bm2, v);
          s_best = kNone;
        } else {
          if (b64 >= next_empty) *a.tot_len = b64 + 1;
          s_best = u32(b64);
          a.f[v] = u32(b64);
          a.tot[cur] -= kv;
          a.tot[b64] += kv;
          a.B[v] = 1;
        }
      }
    }
    __threadfence_block();
    __syncthreads();
    const u32 bst = s_best;
    const bool moved = s_bg > 0.0 && bst != kNone;
    const bool stop = s_bg > 0.0 && bst == kNone;
    if (moved) {
      const u32 sv = a.s[v];
      for (u32 j = tid; j < deg; j += 256) {
        u32 u = g.col[off + j];
        if (a.f[u] != bst) bm_set(a.bm0, a.bm1, a.bm2, u);
        if (a.s[u] == sv) {
          a.K[u] = 1;
          a.K[v] = 1;
        }
      }
    }
    for (u32 q = tid; q < nused; q += 256) {
      u32 h = used[q];
      HK[h] = kNone;
      HV[h] = 0.0;
    }
    __threadfence_block();
    __syncthreads();
    if (tid == 0) s_best = kNone, s_bg = 0.0;
    __syncthreads();
    if (stop) break;
  }
  if (tid == 0) atomicAdd((unsigned long long*)a.pops, (unsigned long long)pops);
}

// ------------------------------------------------------------------ CC split
// aff[s(v)] for v in K; status[0] = some v in K, status[1] = s is not the identity
__global__ void k_aff_mark(const u8* K, const u32* s, u32 n, u8* aff, u32* status) {
  bool any = false, nid = false;
  for (u64 v = gtid(); v < n; v += gstride()) {
    const u32 sv = s[v];
    if (K[v]) {
      aff[sv] = 1;
      any = true;
    }
    nid |= sv != v;
  }
  if (__any(any) && (threadIdx.x & 63) == 0) status[0] = 1;
  if (__any(nid) && (threadIdx.x & 63) == 0) status[1] = 1;
}
__global__ void k_incc(const u32* s, u32 n, const u8* aff, u8* incc) {
  for (u64 v = gtid(); v < n; v += gstride()) incc[v] = aff[s[v]];
}
// also clears the hub-row count of the edge pass that follows
__global__ void k_uf_init(const u32* vcc, u64 m, u32* parent, u32* cnt, u32* hub_cnt) {
  if (gtid() == 0) *hub_cnt = 0;
  for (u64 i = gtid(); i < m; i += gstride()) {
    u32 v = vcc[i];
    parent[v] = v;
    cnt[v] = 0;
  }
}
__device__ __forceinline__ u32 uf_find(u32* parent, u32 x) {
  while (true) {
    u32 p = parent[x];
    if (p == x) return x;
    u32 gp = parent[p];
    if (gp != p) atomicCAS(&parent[x], p, gp);  // path halving
    x = p;
  }
}
__device__ __forceinline__ void uf_union(u32* parent, u32 a, u32 b) {
  while (true) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    u32 hi = max(a, b), lo = min(a, b);
    if (atomicCAS(&parent[hi], hi, lo) == hi) return;
    a = hi;
    b = lo;
  }
}
// edges (v,u) with s(u) = s(v), u > v (and f(u) = f(v) when f is given): union;
// roots end as the component minimum
// A row above kIntraHub entries goes to the hub list (for a kernel that walks
// it in pieces over the grid) instead of the one wavefront that reached it;
// true when listed.  Every lane of the wavefront calls it for the same row.
__device__ __forceinline__ bool wave_list_hub(u32 v, u32 deg, u32* hubs, u32* nhubs, u32 hub_cap) {
  if (!hubs || deg <= kIntraHub) return false;
  u32 slot = 0;
  if ((threadIdx.x & 63) == 0) slot = atomicAdd(nhubs, 1u);
  slot = __shfl(slot, 0, 64);
  if (slot >= hub_cap) return false;
  if ((threadIdx.x & 63) == 0) hubs[slot] = v;
  return true;
}
// piece q of the listed hubs: its row and entry range; false when empty
__device__ __forceinline__ bool hub_piece(const Csr& g, const u32* hubs, u64 q, u32 pieces, u32& v, u64& b,
                                          u64& e) {
  v = hubs[q / pieces];
  b = (q % pieces) * u64(kIntraPiece);
  const u32 deg = g.row_len[v];
  if (b >= deg) return false;
  e = min<u64>(deg, b + kIntraPiece);
  return true;
}
__global__ __launch_bounds__(256) void k_uf_edges(Csr g, const u32* vcc, u64 m, const u32* s, const u32* f,
                                                  u32* parent, u32* hubs, u32* nhubs, u32 hub_cap) {
  const u32 lane = threadIdx.x & 63;
  const u64 wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (u64 q = wave; q < m; q += nwaves) {
    u32 v = vcc[q];
    u64 off = g.row_off[v];
    u32 deg = g.row_len[v];
    if (wave_list_hub(v, deg, hubs, nhubs, hub_cap)) continue;
    u32 sv = s[v];
    const u32 fv = f ? f[v] : 0;
    for (u32 j = lane; j < deg; j += 64) {
      u32 u = g.col[off + j];
      if (u > v && s[u] == sv && (!f || f[u] == fv)) uf_union(parent, v, u);
    }
  }
}
__global__ __launch_bounds__(256) void k_uf_edges_hubs(Csr g, const u32* s, const u32* f, u32* parent,
                                                       const u32* hubs, const u32* nhubs, u32 hub_cap, u32 pieces) {
  const u32 nh = min(*nhubs, hub_cap);
  for (u64 q = blockIdx.x; q < u64(nh) * pieces; q += gridDim.x) {
    u32 v;
    u64 b, e;
    if (!hub_piece(g, hubs, q, pieces, v, b, e)) continue;
    const u64 off = g.row_off[v];
    const u32 sv = s[v], fv = f ? f[v] : 0;
    for (u64 j = b + threadIdx.x; j < e; j += 256) {
      const u32 u = g.col[off + j];
      if (u > v && s[u] == sv && (!f || f[u] == fv)) uf_union(parent, v, u);
    }
  }
}
// cnt[key] += 1 for the lanes that want it, one atomic per distinct key of the
// wavefront (the members of a large component share its root, and atomics on
// one address serialise)
__device__ __forceinline__ void wave_count_keyed(u32* cnt, u32 key, bool want) {
  unsigned long long act = __ballot(want);
  const u32 lane = threadIdx.x & 63;
  while (act) {
    const int leader = __ffsll((long long)act) - 1;
    const u32 k = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(want && key == k) & act;
    if (lane == u32(leader)) atomicAdd(&cnt[k], u32(__popcll(same)));
    act &= ~same;
  }
}
__global__ void k_uf_compress(const u32* vcc, u64 m, u32* parent, u32* cnt) {
  for (u64 base = gtid() - (threadIdx.x & 63); base < m; base += gstride()) {
    const u64 i = base + (threadIdx.x & 63);
    u32 r = 0;
    if (i < m) {
      const u32 v = vcc[i];
      r = uf_find(parent, v);
      parent[v] = r;
    }
    wave_count_keyed(cnt, r, i < m);
  }
}
__global__ void k_zero_best(const u32* vcc, u64 m, const u32* s, u64* best) {
  for (u64 i = gtid(); i < m; i += gstride()) best[s[vcc[i]]] = 0;
}
// keep the LAST largest component (max_by_key, hit_leiden.rs:1317-1322):
// discovery order is ascending minimum vertex, so the key is (size, root)
// the kept component of each subcommunity: the largest, ties to the larger
// root (the last discovered, hit_leiden.rs:1317-1322), or with first_largest
// to the smaller root (the static path's rule, :505-518); roots are the
// components' minimum vertices
__global__ void k_best_comp(const u32* vcc, u64 m, const u32* parent, const u32* cnt, const u32* s, u64* best,
                            bool first_largest) {
  for (u64 i = gtid(); i < m; i += gstride()) {
    u32 v = vcc[i];
    if (parent[v] != v) continue;
    unsigned long long key = (u64(cnt[v]) << 32) | (first_largest ? ~v : v);
    atomicMax((unsigned long long*)&best[s[v]], key);
  }
}
__global__ void k_split_roots(const u32* vcc, u64 m, const u32* parent, const u32* s, const u64* best, u8* flag,
                              bool first_largest) {
  for (u64 i = gtid(); i < m; i += gstride()) {
    u32 v = vcc[i];
    flag[i] = (parent[v] == v) && (u32(best[s[v]] & 0xffffffffu) != (first_largest ? ~v : v));
  }
}
__global__ void k_root_keys(const u32* sel, u64 c, const u32* vcc, const u32* s, u64* keys, u32* vals) {
  for (u64 r = gtid(); r < c; r += gstride()) {
    u32 v = vcc[sel[r]];
    keys[r] = (u64(s[v]) << 32) | v;
    vals[r] = v;
  }
}
__global__ void k_newids(const u32* roots_sorted, u64 c, u32 base, u32* newid) {
  for (u64 r = gtid(); r < c; r += gstride()) newid[roots_sorted[r]] = base + u32(r);
}
__global__ void k_split_assign(const u32* vcc, u64 m, const u32* parent, const u64* best, const u32* newid, u32* s,
                               u8* R, const double* k, SubRef sub, bool first_largest) {
  for (u64 i = gtid(); i < m; i += gstride()) {
    u32 v = vcc[i];
    u32 r = parent[v];
    u32 sv = s[v];
    if (u32(best[sv] & 0xffffffffu) != (first_largest ? ~r : r)) {
      const u32 to = newid[r];
      s[v] = to;
      R[v] = 1;
      if (sub.size) {
        atomicSub(&sub.size[sv], 1u);
        atomicAdd(&sub.size[to], 1u);
        atomicAdd(&sub.deg[sv], -k[v]);
        atomicAdd(&sub.deg[to], k[v]);
      }
    }
  }
}
__global__ void k_sub_add(const u32* s, u32 n0, u32 n1, u32* size) {
  for (u64 v = n0 + gtid(); v < n1; v += gstride()) atomicAdd(&size[s[v]], 1u);
}

// ------------------------------------------------------------------ refinement (throughput)
__global__ void k_sizes(const u32* s, u32 n, u32* sizes) {
  for (u64 v = gtid(); v < n; v += gstride()) atomicAdd(&sizes[s[v]], 1u);
}
// v in R and alone in its subcommunity
// the refinement's per-subcommunity arrays cleared in one launch: candidate
// flags, parent members (none), w(S, C), and the candidate marks' hub-row count
__global__ void k_refine_clear(u8* is_cand, u32* pidx, double* wS, u64 sn, u32* hub_cnt) {
  if (gtid() == 0) *hub_cnt = 0;
  for (u64 i = gtid(); i < sn; i += gstride()) {
    is_cand[i] = 0;
    pidx[i] = kNone;
    wS[i] = 0.0;
  }
}
// the label array of cc_split cleared, and its three status words
__global__ void k_cc_clear(u8* aff, u64 n, u32* status) {
  if (gtid() < 3) status[gtid()] = 0;
  for (u64 i = gtid(); i < n; i += gstride()) aff[i] = 0;
}
__global__ void k_single(const u8* R, u32 n, const u32* s, const u32* sizes, u8* single) {
  for (u64 v = gtid(); v < n; v += gstride()) single[v] = R[v] && sizes[s[v]] == 1;
}
// candidate subcommunities of singleton R vertices (hit_leiden.rs:1569-1586)
__global__ __launch_bounds__(256) void k_cand_mark(Csr g, const u32* sl, u64 ns, const u32* f, const u32* s,
                                                    u8* is_cand, u32* hubs, u32* nhubs, u32 hub_cap) {
  const u32 lane = threadIdx.x & 63;
  const u64 wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (u64 q = wave; q < ns; q += nwaves) {
    u32 v = sl[q];
    u64 off = g.row_off[v];
    u32 deg = g.row_len[v];
    if (wave_list_hub(v, deg, hubs, nhubs, hub_cap)) continue;
    u32 fv = f[v], sv = s[v];
    for (u32 j = lane; j < deg; j += 64) {
      u32 u = g.col[off + j];
      if (f[u] == fv && s[u] != sv) is_cand[s[u]] = 1;
    }
  }
}
__global__ __launch_bounds__(256) void k_cand_mark_hubs(Csr g, const u32* f, const u32* s, u8* is_cand,
                                                        const u32* hubs, const u32* nhubs, u32 hub_cap,
                                                        u32 pieces) {
  const u32 nh = min(*nhubs, hub_cap);
  for (u64 q = blockIdx.x; q < u64(nh) * pieces; q += gridDim.x) {
    u32 v;
    u64 b, e;
    if (!hub_piece(g, hubs, q, pieces, v, b, e)) continue;
    const u64 off = g.row_off[v];
    const u32 fv = f[v], sv = s[v];
    for (u64 j = b + threadIdx.x; j < e; j += 256) {
      const u32 u = g.col[off + j];
      if (f[u] == fv && s[u] != sv) is_cand[s[u]] = 1;
    }
  }
}
// also clears the hub-row count of k_ws (the candidate marks' hub rows are done)
__global__ void k_parent_idx(const u32* s, u32 n, const u8* is_cand, u32* pidx, u32* hub_cnt) {
  if (gtid() == 0) *hub_cnt = 0;
  for (u64 v = gtid(); v < n; v += gstride())
    if (is_cand[s[v]]) atomicMin(&pidx[s[v]], u32(v));
}
// w(S, C) for the candidate subcommunities S (hit_leiden.rs:1617-1629: the
// weight from S's members to their parent community C, a self-loop twice).
// Each wavefront takes 64 consecutive vertices.  Their rows (up to kWsWide
// entries) are walked as one flat sequence, 64 entries per step (the row of
// entry t by a binary search over the lanes' degree prefix); a segmented scan
// over the lanes sums each row's part of the step, and the last lane of the
// segment adds it to the row's total in LDS (one writer per row and step, so
// no atomics).  Longer rows are summed by the whole wavefront, one after
// another.  The 64 vertex totals are then combined per subcommunity in the
// wavefront (one reduction per distinct S among them) before one keyed add per
// S into the workgroup's LDS table, flushed with one device atomic per S and
// pass: the members of a big subcommunity would otherwise serialise on one
// address, in LDS as in HBM.  Rows above kIntraHub entries are listed for
// k_ws_hubs (pieces over the grid: one wavefront would walk a hub of 10^6
// entries for tens of ms while the rest of the grid idles); when the list is
// full the wavefront walks them itself.
constexpr u32 kWsTable = 2048, kWsWide = 1024;
__global__ __launch_bounds__(256) void k_ws(Csr g, u32 n, const u32* f, const u32* s, const u8* is_cand,
                                             const u32* pidx, double* wS, u32* hubs, u32* nhubs, u32 hub_cap,
                                             u64 hub_max) {
  __shared__ u32 tk[kWsTable];
  __shared__ double tv[kWsTable];
  __shared__ double rowacc[256];
  const u32 lane = threadIdx.x & 63;
  double* racc = rowacc + (threadIdx.x & ~63u);
  for (u64 base0 = u64(blockIdx.x) * 256; base0 < n; base0 += u64(gridDim.x) * 256) {
    for (u32 i = threadIdx.x; i < kWsTable; i += 256) {
      tk[i] = kNone;
      tv[i] = 0.0;
    }
    rowacc[threadIdx.x] = 0.0;
    __syncthreads();
    const u64 base = base0 + (threadIdx.x & ~63u);
    const u64 v = base + lane;
    u32 sv = kNone, C = 0, deg = 0;
    u64 off = 0;
    if (v < n) {
      const u32 x = s[v];
      if (is_cand[x]) {
        sv = x;
        C = f[pidx[x]];
        off = g.row_off[v];
        deg = g.row_len[v];
      }
    }
    if (deg > kIntraHub && deg <= hub_max) {
      const u32 slot = atomicAdd(nhubs, 1u);
      if (slot < hub_cap) {
        hubs[slot] = u32(v);
        deg = 0;  // k_ws_hubs adds this row
      }
    }
    const bool wide = deg > kWsWide;
    const u32 d0 = wide ? 0u : deg;
    u32 x = d0;
    for (u32 o = 1; o < 64; o <<= 1) {
      const u32 y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    const u32 total = __shfl(x, 63, 64), pre = x - d0;
    for (u32 t0 = 0; t0 < total; t0 += 64) {
      const u32 t = t0 + lane;
      u32 r = 0;
      for (u32 step = 32; step > 0; step >>= 1) {
        const u32 pc = __shfl(pre, r + step, 64);
        if (pc <= t) r += step;
      }
      const u64 ro = __shfl((unsigned long long)off, r, 64);
      const u32 rp = __shfl(pre, r, 64), rC = __shfl(C, r, 64);
      double val = 0.0;
      if (t < total) {
        const u64 e = ro + (t - rp);
        const u32 u = g.col[e];
        if (f[u] == rC) {
          const double w = g.w[e];
          val = (u == base + r) ? 2.0 * w : w;
        }
      }
      // inclusive segmented sum over lanes of the same row (rows are
      // contiguous in lane order)
      for (u32 o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(val, o, 64);
        const u32 ry = __shfl_up(r, o, 64);
        if (lane >= o && ry == r) val += y;
      }
      const u32 rn = __shfl_down(r, 1, 64);
      const bool last = t < total && (lane == 63 || t + 1 >= total || rn != r);
      if (last) racc[r] += val;
    }
    unsigned long long mask = __ballot(wide);
    while (mask) {
      const u32 i = u32(__ffsll((long long)mask) - 1);
      mask &= mask - 1;
      const u64 bv = base + i;
      const u64 boff = __shfl((unsigned long long)off, i, 64);
      const u32 bdeg = __shfl(deg, i, 64), bC = __shfl(C, i, 64);
      double acc = 0.0;
      for (u32 j = lane; j < bdeg; j += 64) {
        const u32 u = g.col[boff + j];
        if (f[u] == bC) {
          const double w = g.w[boff + j];
          acc += (u == bv) ? 2.0 * w : w;
        }
      }
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) racc[i] = acc;
    }
    __threadfence_block();
    // the wavefront's vertex totals, combined per subcommunity
    const double mine = racc[lane];
    bool pending = sv != kNone && mine != 0.0;
    unsigned long long pm = __ballot(pending);
    while (pm) {
      const u32 lead = u32(__ffsll((long long)pm) - 1);
      const u32 key = __shfl(sv, lead, 64);
      const bool member = pending && sv == key;
      double y = member ? mine : 0.0;
      for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o, 64);
      if (lane == lead) lds_add_keyed(tk, tv, kWsTable, key, y);
      pending = pending && !member;
      pm = __ballot(pending);
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < kWsTable; i += 256)
      if (tk[i] != kNone && tv[i] != 0.0) atomicAdd(&wS[tk[i]], tv[i]);
    __syncthreads();
  }
}
// the listed hub rows of k_ws in pieces of kIntraPiece entries, one
// workgroup per piece and one device atomic per piece
__global__ __launch_bounds__(256) void k_ws_hubs(Csr g, const u32* f, const u32* s, const u32* pidx,
                                                 const u32* hubs, const u32* nhubs, u32 hub_cap, u32 pieces,
                                                 double* wS) {
  const u32 nh = min(*nhubs, hub_cap);
  for (u64 q = blockIdx.x; q < u64(nh) * pieces; q += gridDim.x) {
    const u32 v = hubs[q / pieces];
    const u64 b = (q % pieces) * u64(kIntraPiece);
    const u32 deg = g.row_len[v];
    if (b >= deg) continue;
    const u64 end = min<u64>(deg, b + kIntraPiece), off = g.row_off[v];
    const u32 S = s[v], C = f[pidx[S]];
    double local = 0.0;
    for (u64 j = b + threadIdx.x; j < end; j += 256) {
      const u32 u = g.col[off + j];
      if (f[u] == C) {
        const double w = g.w[off + j];
        local += (u == v) ? 2.0 * w : w;
      }
    }
    const double t = block_sum(local);
    if (threadIdx.x == 0 && t != 0.0) atomicAdd(&wS[S], t);
  }
}
__global__ void k_blocked(u64 sn, const u8* is_cand, const double* dS, const double* wS, const u32* pidx,
                          const u32* f, const double* dpar, const double* twice_m, const u32* gid, double gamma,
                          u8* blocked) {
  for (u64 c = gtid(); c < sn; c += gstride()) {
    u8 b = 0;
    if (is_cand[c]) {
      double ds = dS[c];
      if (ds > 0.0) {
        const u32 x = pidx[c];  // a member of subcommunity c
        const double t = gid ? twice_m[gid[x]] : twice_m[0];
        u32 P = f[x];
        b = remove_gain(wS[c], t, gamma, ds, dpar[P]) > kEpsTFilter;
      }
    }
    blocked[c] = b;
  }
}
__global__ void k_gather_list(const u32* sel, u64 c, const u32* from, u32* out) {
  for (u64 i = gtid(); i < c; i += gstride()) out[i] = from[sel[i]];
}

// ------------------------------------------------------------------ refinement (deterministic)
struct DetRefArgs {
  Csr g;
  const u32* f;
  u32* s;
  const u32* order;  // R sorted by (k asc, id asc)
  u64 nr;
  u32* sizes;
  double* dS;
  const double* dpar;
  u32* mhead;
  u32* mnext;
  double gamma;
  u32* gk;
  double* gv;
  u32* gused;
  u32 gcap;
};
// hit_leiden.rs:1394-1511: singleton merges in order, T-filter per candidate
// over the current members of S (ascending), sequential per-key sums
__global__ __launch_bounds__(256) void k_det_refine(DetRefArgs a) {
  __shared__ u32 lk[DET_LCAP];
  __shared__ double lv[DET_LCAP];
  __shared__ u32 lused[DET_LCAP];
  __shared__ u32 cl[256];
  __shared__ double cw[256];
  __shared__ u8 cs[256];
  __shared__ u32 s_nused, s_best;
  __shared__ double s_wcur, s_bg;
  const u32 tid = threadIdx.x;
  for (u32 q = tid; q < DET_LCAP; q += 256) {
    lk[q] = kNone;
    lv[q] = 0.0;
  }
  __syncthreads();
  const Csr& g = a.g;
  const double tt = *g.twice_m;
  for (u64 it = 0; it < a.nr; ++it) {
    const u32 v = a.order[it];
    const u32 sc = a.s[v];
    if (a.sizes[sc] != 1) continue;  // uniform across the block
    const u32 C = a.f[v];
    const double kv = g.k[v];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const bool in_lds = deg <= DET_LCAP / 2;
    u32* HK = in_lds ? lk : a.gk;
    double* HV = in_lds ? lv : a.gv;
    u32* used = in_lds ? lused : a.gused;
    const u32 cap = in_lds ? DET_LCAP : a.gcap;
    if (tid == 0) s_nused = 0;
    __syncthreads();
    for (u32 base = 0; base < deg; base += 256) {
      u32 j = base + tid;
      if (j < deg) {
        u32 u = g.col[off + j];
        bool in = a.f[u] == C;
        cl[tid] = in ? a.s[u] : kNone;
        cw[tid] = g.w[off + j];
        cs[tid] = (u == v);
      }
      __syncthreads();
      if (tid == 0) {
        u32 m = min(256u, deg - base);
        u32 nu = s_nused;
        for (u32 q = 0; q < m; ++q) {
          if (cl[q] == kNone) continue;
          u32 h = seq_insert(HK, cap, cl[q], used, nu);
          HV[h] += cw[q];
          if (cs[q]) HV[h] += cw[q];
        }
        s_nused = nu;
      }
      __syncthreads();
    }
    if (tid == 0) s_wcur = table_lookup(HK, HV, cap, sc);
    __syncthreads();
    const double wcur = s_wcur;
    const u32 nused = s_nused;
    const double dp = a.dpar[C];
    const double dcur = a.dS[sc];
    double bg = 0.0;
    u32 best = kNone;
    for (u32 q = tid; q < nused; q += 256) {
      u32 h = used[q];
      u32 S = HK[h];
      if (S == sc) continue;
      double ds = a.dS[S];
      if (ds > 0.0) {
        double ws = 0.0;
        for (u32 i = a.mhead[S]; i != kNone; i = a.mnext[i]) {
          u64 o2 = g.row_off[i];
          u32 d2 = g.row_len[i];
          for (u32 j = 0; j < d2; ++j) {
            u32 x = g.col[o2 + j];
            if (a.f[x] == C) {
              double w = g.w[o2 + j];
              ws += w;
              if (x == i) ws += w;
            }
          }
        }
        if (remove_gain(ws, tt, a.gamma, ds, dp) > kEpsTFilter) continue;
      }
      double gn = move_gain(HV[h], wcur, tt, a.gamma, kv, dcur, ds);
      if (gn > 0.0) better(bg, best, gn, S);
    }
    block_argmax(bg, best);
    if (tid == 0 && bg > 0.0) {
      a.s[v] = best;
      a.sizes[sc] -= 1;
      a.sizes[best] += 1;
      a.dS[sc] -= kv;
      a.dS[best] += kv;
      a.mhead[sc] = kNone;
      u32 prev = kNone, x = a.mhead[best];
      while (x != kNone && x < v) {
        prev = x;
        x = a.mnext[x];
      }
      a.mnext[v] = x;
      if (prev == kNone)
        a.mhead[best] = v;
      else
        a.mnext[prev] = v;
    }
    for (u32 q = tid; q < nused; q += 256) {
      u32 h = used[q];
      HK[h] = kNone;
      HV[h] = 0.0;
    }
    __threadfence_block();
    __syncthreads();
  }
}
__global__ void k_member_lists(u64 nseg, const u32* starts, u64 n, const u32* keys_s, const u32* vals_s, u32* mhead,
                               u32* mnext) {
  for (u64 j = gtid(); j < nseg; j += gstride()) {
    u64 s = starts[j], e = (j + 1 < nseg) ? starts[j + 1] : n;
    mhead[keys_s[s]] = vals_s[s];
    for (u64 t = s; t < e; ++t) mnext[vals_s[t]] = (t + 1 < e) ? vals_s[t + 1] : kNone;
  }
}
__global__ void k_degree_keys(const u32* rl, u64 nr, const double* k, u64* keys) {
  for (u64 i = gtid(); i < nr; i += gstride()) keys[i] = (u64)__double_as_longlong(k[rl[i]]);
}

// ------------------------------------------------------------------ aggregation
__global__ void k_lift(const Rec* d, u64 k, const u32* s_pre, Rec* H) {
  for (u64 i = gtid(); i < k; i += gstride()) {
    Rec r = d[i];
    H[i] = Rec{s_pre[r.u], s_pre[r.v], r.w};
  }
}
// hit_leiden.rs:1659-1692 with an R-R edge emitted once (v_j not in R or i < j)
// and the self-loop record weight transferred
__device__ __forceinline__ bool agg_q(u32 v, u32 u, const u8* R) { return u != v && (!R[u] || v < u); }

__global__ __launch_bounds__(256) void k_agg_count(Csr g, const u32* rl, u64 nr, const u8* R, u64* cnt) {
  const u32 lane = threadIdx.x & 63;
  const u64 wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (u64 q = wave; q < nr; q += nwaves) {
    u32 v = rl[q];
    u64 off = g.row_off[v];
    u32 deg = g.row_len[v];
    u32 c = 0;
    u32 self = 0;
    for (u32 j = lane; j < deg; j += 64) {
      u32 u = g.col[off + j];
      c += agg_q(v, u, R);
      if (u == v && fabs(g.w[off + j]) > kEpsSelf) self = 1;
    }
    for (int o = 32; o > 0; o >>= 1) {
      c += __shfl_xor(c, o, 64);
      self |= __shfl_xor(self, o, 64);
    }
    if (lane == 0) cnt[q] = 2ull * c + (self ? 2 : 0);
  }
}
__global__ __launch_bounds__(256) void k_agg_emit(Csr g, const u32* rl, u64 nr, const u8* R, const u32* s_pre,
                                                   const u32* s_cur, const u64* offs, u64 base, Rec* H) {
  const u32 lane = threadIdx.x & 63;
  const u64 wave = gtid() >> 6, nwaves = gstride() >> 6;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (u64 q = wave; q < nr; q += nwaves) {
    u32 v = rl[q];
    u64 off = g.row_off[v];
    u32 deg = g.row_len[v];
    u64 pos = base + offs[q];
    const u32 pv = s_pre[v], cv = s_cur[v];
    double self = 0.0;
    for (u32 j0 = 0; j0 < deg; j0 += 64) {
      u32 j = j0 + lane;
      bool qual = false;
      u32 u = 0;
      double w = 0.0;
      if (j < deg) {
        u = g.col[off + j];
        w = g.w[off + j];
        qual = agg_q(v, u, R);
        if (u == v) self = w;
      }
      unsigned long long m = __ballot(qual);
      if (qual) {
        u64 p = pos + 2ull * __popcll(m & lt);
        H[p] = Rec{pv, s_pre[u], -w};
        H[p + 1] = Rec{cv, s_cur[u], w};
      }
      pos += 2ull * __popcll(m);
    }
    for (int o = 32; o > 0; o >>= 1) {
      double os = __shfl_xor(self, o, 64);
      if (os != 0.0) self = os;
    }
    if (lane == 0 && fabs(self) > kEpsSelf) {
      H[pos] = Rec{pv, pv, -self};
      H[pos + 1] = Rec{cv, cv, self};
    }
  }
}
__global__ void k_spre_update(const u32* rl, u64 nr, const u32* s_cur, u32* s_pre) {
  for (u64 i = gtid(); i < nr; i += gstride()) s_pre[rl[i]] = s_cur[rl[i]];
}
__global__ void k_hkeys(const Rec* H, u64 n, u64* keys, u32* idx) {
  for (u64 i = gtid(); i < n; i += gstride()) {
    Rec r = H[i];
    u32 a = min(r.u, r.v), b = max(r.u, r.v);
    keys[i] = (u64(a) << 32) | b;
    idx[i] = u32(i);
  }
}
// Compress (hit_leiden.rs:1701-1714): per key in emission order, |w| > 1e-9 kept
// nruns from the device (the run selection's count); keep[nruns..n) is
// cleared so that the survivors can be selected over all n flags
__global__ void k_compress(const unsigned long long* nruns_dev, const u32* runs, u64 n, const u64* keys,
                           const u32* idx, const Rec* H, Rec* out, u8* keep) {
  const u64 nruns = *nruns_dev;
  for (u64 r = nruns + gtid(); r < n; r += gstride()) keep[r] = 0;
  for (u64 r = gtid(); r < nruns; r += gstride()) {
    u64 s = runs[r], e = (r + 1 < nruns) ? runs[r + 1] : n;
    double acc = 0.0;
    for (u64 t = s; t < e; ++t) acc += H[idx[t]].w;
    u64 key = keys[s];
    out[r] = Rec{u32(key >> 32), u32(key & 0xffffffffu), acc};
    keep[r] = fabs(acc) > kEpsCompress;
  }
}
__global__ void k_gather_rec(const u32* sel, u64 c, const Rec* in, Rec* out) {
  for (u64 i = gtid(); i < c; i += gstride()) out[i] = in[sel[i]];
}

// ------------------------------------------------------------------ def_update / modularity
// tot[key] += x for the lanes that want it: the lanes of the first few
// distinct keys of the wavefront are summed first (one atomic per key: a pull
// moves whole communities, and atomics on one address serialise), the rest
// add their own
__device__ __forceinline__ void wave_add_keyed(double* tot, u32 key, double x, bool want) {
  const u32 lane = threadIdx.x & 63;
  unsigned long long act = __ballot(want);
  for (int it = 0; it < 4 && act; ++it) {
    const int leader = __ffsll((long long)act) - 1;
    const u32 kk = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(want && key == kk) & act;
    double y = ((same >> lane) & 1ull) ? x : 0.0;
    for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o, 64);
    if (lane == u32(leader)) atomicAdd(&tot[kk], y);
    act &= ~same;
  }
  if ((act >> lane) & 1ull) atomicAdd(&tot[key], x);
}
__global__ void k_def_pull(const u8* B, u32* fp, const u32* fq, const u32* sp, u32 n, double* tot, const double* k) {
  for (u64 base = gtid() - (threadIdx.x & 63); base < n; base += gstride()) {
    const u64 v = base + (threadIdx.x & 63);
    u32 old = 0, nw = 0;
    double kv = 0.0;
    bool ch = false;
    if (v < n && (!B || B[v])) {
      old = fp[v];
      nw = fq[sp[v]];
      fp[v] = nw;
      if (tot && old != nw) {
        kv = k[v];
        ch = kv != 0.0;
      }
    }
    if (tot) {  // keep the maintained community totals current
      wave_add_keyed(tot, old, -kv, ch);
      wave_add_keyed(tot, nw, kv, ch);
    }
  }
}
__global__ void k_def_push(const u8* Bp, u8* Bq, const u32* sq, u32 nq, u64 np) {
  for (u64 u = gtid(); u < nq; u += gstride()) {
    u32 x = sq[u];
    if (x < np && Bp[x]) Bq[u] = 1;
  }
}
__global__ void k_gather_u32(u32* cur, const u32* map, u32 n) {
  for (u64 v = gtid(); v < n; v += gstride()) cur[v] = map[cur[v]];
}
// Intra-community weight (every edge from both ends, a self-loop's entry
// twice; compute_modularity, hit_leiden.rs:121-161).  Every edge is stored in
// both of its rows, so 2w over the entries with u >= v (a self-loop is the one
// entry u = v) is the same sum with half the label gathers.  Each wavefront takes 64
// consecutive vertices and walks their rows as one concatenated sequence, 64
// consecutive entries per step (the row of entry t is found by a binary search
// over the lanes' degree prefix), so the lanes load neighbouring addresses and
// none idles on a short row.  Rows above kIntraHub entries are listed for
// k_intra_hubs; when the list is full the wavefront walks them itself.
__global__ __launch_bounds__(256) void k_intra(Csr g, const u32* lab, double* acc, u32* hubs, u32* nhubs,
                                               u32 hub_cap, u64 hub_max) {
  double local = 0.0;
  const u32 lane = threadIdx.x & 63;
  for (u64 base = gtid() - lane; base < g.n; base += gstride()) {
    const u64 v = base + lane;
    u64 off = 0;
    u32 deg = 0, lv = 0;
    if (v < g.n) {
      off = g.row_off[v];
      deg = g.row_len[v];
      lv = lab[v];
    }
    bool wide = deg > kIntraHub;
    if (wide && deg <= hub_max) {  // hub_max: the rows k_intra_hubs covers in full
      const u32 slot = atomicAdd(nhubs, 1u);
      if (slot < hub_cap) {
        hubs[slot] = u32(v);
        wide = false;
        deg = 0;
      }
    }
    const u32 d0 = wide ? 0u : deg;
    u32 x = d0;  // inclusive prefix of d0 over the lanes
    for (u32 o = 1; o < 64; o <<= 1) {
      const u32 y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    const u32 total = __shfl(x, 63, 64), pre = x - d0;
    for (u32 t0 = 0; t0 < total; t0 += 64) {
      const u32 t = t0 + lane;
      u32 r = 0;  // the last lane whose prefix is <= t
      for (u32 step = 32; step > 0; step >>= 1) {
        const u32 pc = __shfl(pre, r + step, 64);
        if (pc <= t) r += step;
      }
      const u64 ro = __shfl((unsigned long long)off, r, 64);
      const u32 rp = __shfl(pre, r, 64), rl = __shfl(lv, r, 64);
      if (t < total) {
        const u64 e = ro + (t - rp);
        const u32 u = g.col[e];
        if (u >= base + r && lab[u] == rl) local += 2.0 * g.w[e];
      }
    }
    unsigned long long m = __ballot(wide);
    while (m) {
      const int src = __ffsll((long long)m) - 1;
      m &= m - 1;
      const u64 bv = __shfl((unsigned long long)v, src, 64);
      const u64 boff = __shfl((unsigned long long)off, src, 64);
      const u32 bdeg = __shfl(deg, src, 64), blv = __shfl(lv, src, 64);
      for (u32 j = lane; j < bdeg; j += 64) {
        const u32 u = g.col[boff + j];
        if (u >= bv && lab[u] == blv) local += 2.0 * g.w[boff + j];
      }
    }
  }
  double t = block_sum(local);
  if (threadIdx.x == 0 && t != 0.0) atomicAdd(acc, t);
}
// the listed hub rows in pieces of kIntraPiece entries, one workgroup per piece
__global__ __launch_bounds__(256) void k_intra_hubs(Csr g, const u32* lab, const u32* hubs, const u32* nhubs,
                                                    u32 hub_cap, u32 pieces, double* acc) {
  const u32 nh = min(*nhubs, hub_cap);
  double local = 0.0;
  for (u64 q = blockIdx.x; q < u64(nh) * pieces; q += gridDim.x) {
    const u32 v = hubs[q / pieces];
    const u64 b = (q % pieces) * u64(kIntraPiece);
    const u32 deg = g.row_len[v];
    if (b >= deg) continue;
    const u64 end = min<u64>(deg, b + kIntraPiece), off = g.row_off[v];
    const u32 lv = lab[v];
    for (u64 j = b + threadIdx.x; j < end; j += 256) {
      const u32 u = g.col[off + j];
      if (u >= v && lab[u] == lv) local += 2.0 * g.w[off + j];
    }
  }
  double t = block_sum(local);
  if (threadIdx.x == 0 && t != 0.0) atomicAdd(acc, t);
}
__global__ void k_sq_sum(const double* sig, u64 n, const double* twice_m, double* acc) {
  const double t = *twice_m;
  double local = 0.0;
  for (u64 i = gtid(); i < n; i += gstride()) {
    double x = sig[i];
    local += x * x / t;
  }
  double s = block_sum(local);
  if (threadIdx.x == 0 && s != 0.0) atomicAdd(acc, s);
}
__global__ void k_first_seen(const u32* lab, u32 n, u32* first) {
  for (u64 v = gtid(); v < n; v += gstride()) atomicMin(&first[lab[v]], u32(v));
}
__global__ void k_first_flags(const u32* lab, u32 n, const u32* first, u64* flag) {
  for (u64 v = gtid(); v < n; v += gstride()) flag[v] = first[lab[v]] == v;
}
__global__ void k_relabel(u32* lab, u32 n, const u32* first, const u64* pos) {
  for (u64 v = gtid(); v < n; v += gstride()) lab[v] = u32(pos[first[lab[v]]]);
}

// ------------------------------------------------------------------ partition comparison
__global__ void k_pair_keys(const u32* a, const u32* b, u64 n, u64* keys, u32* sa, u32* sb) {
  for (u64 i = gtid(); i < n; i += gstride()) {
    keys[i] = (u64(a[i]) << 32) | b[i];
    atomicAdd(&sa[a[i]], 1u);
    atomicAdd(&sb[b[i]], 1u);
  }
}
// one thread per (a, b) run of the sorted pairs: mutual-information term,
// best overlap per label, and the overlap of the two named communities
__global__ void k_run_terms(const u64* keys, const u32* starts, u64 nruns, u64 n, const u32* sa, const u32* sb,
                            u32* besta, u32* bestb, u32 la, u32 lb, double* mi, u32* inter) {
  double local = 0.0;
  const double nf = double(n);
  for (u64 r = gtid(); r < nruns; r += gstride()) {
    const u64 st = starts[r], en = r + 1 < nruns ? starts[r + 1] : n;
    const u32 c = u32(en - st);
    const u64 key = keys[st];
    const u32 x = u32(key >> 32), y = u32(key);
    const double phs = c / nf, ph = sa[x] / nf, ps = sb[y] / nf;
    local += phs * log(phs / (ph * ps));
    atomicMax(&besta[x], c);
    atomicMax(&bestb[y], c);
    if (x == la && y == lb) *inter = c;
  }
  const double t = block_sum(local);
  if (threadIdx.x == 0 && t != 0.0) atomicAdd(mi, t);
}
// entropy of the label sizes and the sum of the best overlaps
__global__ void k_label_terms(const u32* size, const u32* best, u64 L, u64 n, double* ent, unsigned long long* pur) {
  double h = 0.0;
  unsigned long long b = 0;
  for (u64 c = gtid(); c < L; c += gstride()) {
    const u32 x = size[c];
    if (x) {
      const double p = double(x) / double(n);
      h -= p * log(p);
      b += best[c];
    }
  }
  const double t = block_sum(h);
  for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o, 64);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(pur, b);
  if (threadIdx.x == 0 && t != 0.0) atomicAdd(ent, t);
}

// ------------------------------------------------------------------ several graphs per engine
// Vertex ranges of the graphs are contiguous at level 0; gid maps every vertex
// of a level to its graph (G = no graph: an unused id of the level).
__global__ void k_gid_from_offsets(const u32* offs, u32 G, u32 n, u32* gid) {
  for (u64 v = gtid(); v < n; v += gstride()) {
    u32 lo = 0, hi = G;  // offs[lo] <= v < offs[hi]
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (offs[mid] <= v) lo = mid;
      else hi = mid;
    }
    gid[v] = lo;
  }
}
__global__ void k_graph_counts(const Rec* d, u64 k, const u32* gid, u32* cnt) {
  // a graph's records are contiguous: one atomic per graph and wavefront
  for (u64 base = gtid() - (threadIdx.x & 63); base < k; base += gstride()) {
    const u64 i = base + (threadIdx.x & 63);
    wave_count_keyed(cnt, i < k ? gid[d[i].u] : 0u, i < k);
  }
}
// vertices of a list per graph (the lists hold a graph's vertices mostly
// together: one atomic per graph and wavefront)
__global__ void k_list_graph_counts(const u32* list, u64 c, const u32* gid, u32* cnt) {
  for (u64 base = gtid() - (threadIdx.x & 63); base < c; base += gstride()) {
    const u64 i = base + (threadIdx.x & 63);
    wave_count_keyed(cnt, i < c ? gid[list[i]] : 0u, i < c);
  }
}
__global__ void k_part_from_counts(const u32* cnt, u32 G, u8* part) {
  for (u64 g = gtid(); g <= G; g += gstride()) part[g] = g < G && cnt[g] > 0;
}
__global__ void k_graph_nonident(const u32* s, const u32* gid, u32 n, u8* nid) {
  for (u64 v = gtid(); v < n; v += gstride())
    if (s[v] != v) {
      const u32 g = gid[v];
      if (!nid[g]) nid[g] = 1;
    }
}
__global__ void k_fill_R_graphs(const u32* gid, u32 n, const u8* part, const u8* nid, u8* R) {
  for (u64 v = gtid(); v < n; v += gstride()) {
    const u32 g = gid[v];
    if (part[g] && !nid[g]) R[v] = 1;
  }
}
// the graph of each next-level vertex: the graph of a member (every member of
// a subcommunity lies in one graph, so concurrent writes agree); vertices no
// member ever reached (graph G) write nothing
__global__ void k_scatter_gid(const u32* s, const u32* gid_p, u32 n, u32 G, u32* gid_q) {
  for (u64 v = gtid(); v < n; v += gstride()) {
    const u32 g = gid_p[v];
    if (g < G) gid_q[s[v]] = g;
  }
}
// Per-graph modularity terms: sig[c] and the graph of community c, then the
// intra-community weight and the sum of sig^2 per graph.  Each thread walks a
// contiguous chunk (vertices of a graph are mostly adjacent) and adds to the
// device totals when the graph changes.
__global__ void k_sig_graph(const u32* lab, const double* k, const u32* gid, u32 n, double* sig, u32* cg) {
  for (u64 v = gtid(); v < n; v += gstride()) {
    const u32 c = lab[v];
    atomicAdd(&sig[c], k[v]);
    cg[c] = gid[v];
  }
}
__global__ void k_intra_graphs(Csr g, const u32* lab, u64 chunk, double* intra) {
  const u64 t = gtid();
  const u64 v0 = t * chunk, v1 = v0 + chunk < g.n ? v0 + chunk : g.n;
  u32 cur = kNone;
  double acc = 0.0;
  for (u64 v = v0; v < v1; ++v) {
    const u32 gv = g.gid[v];
    if (gv != cur) {
      if (acc != 0.0) atomicAdd(&intra[cur], acc);
      cur = gv;
      acc = 0.0;
    }
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v], lv = lab[v];
    for (u32 j = 0; j < deg; ++j) {
      const u32 u = g.col[off + j];
      if (u >= v && lab[u] == lv) acc += 2.0 * g.w[off + j];
    }
  }
  if (acc != 0.0) atomicAdd(&intra[cur], acc);
}
__global__ void k_sq_graphs(const double* sig, const u32* cg, u64 L, u64 chunk, double* sq) {
  const u64 t = gtid();
  const u64 c0 = t * chunk, c1 = c0 + chunk < L ? c0 + chunk : L;
  u32 cur = kNone;
  double acc = 0.0;
  for (u64 c = c0; c < c1; ++c) {
    const double x = sig[c];
    if (x == 0.0) continue;
    const u32 gc = cg[c];
    if (gc != cur) {
      if (acc != 0.0) atomicAdd(&sq[cur], acc);
      cur = gc;
      acc = 0.0;
    }
    acc += x * x;
  }
  if (acc != 0.0) atomicAdd(&sq[cur], acc);
}


// ------------------------------------------------------------------ static path (run / multilevel_leiden)
constexpr u32 SLCAP = 4096;  // LDS hash slots of the static sweeps

// Ordered append of the flagged lanes of a 256-thread block: returns each
// flagged thread's position among the flagged ones (thread order) and the
// block's count in *total (all threads must call, in step).
__device__ __forceinline__ u32 block_rank(bool want, u32* total) {
  __shared__ u32 wc[4];
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long m = __ballot(want);
  if (lane == 0) wc[wv] = u32(__popcll(m));
  __syncthreads();
  u32 before = 0, all = 0;
  for (u32 i = 0; i < 4; ++i) {
    if (i < wv) before += wc[i];
    all += wc[i];
  }
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  *total = all;
  const u32 r = before + u32(__popcll(m & below));
  __syncthreads();
  return r;
}

struct StaticMoveArgs {
  Csr g;
  u32* f;
  double* cdeg;  // community degrees, indexed by label (labels < n)
  u32* queue;    // ring of n slots; queue[0..qlen) holds the shuffled active vertices
  u32 qlen;
  u8* inq;
  double gamma;
  u32* gk;  // HBM table for rows above SLCAP / 2 entries
  double* gv;
  u32* gused;
  u32 gcap;
  unsigned long long* moves;
};

// single_level_movement's queue loop (hit_leiden.rs:856-928) in one
// workgroup: pop the front, w(v, C) summed in row order without self-loops,
// argmax over neighbour communities in ascending id with strict '>' (no empty
// community), move when positive, then append the neighbours outside the new
// community that are not queued, in row order.
__global__ __launch_bounds__(256) void k_static_move(StaticMoveArgs a) {
  __shared__ u32 lk[SLCAP];
  __shared__ double lv[SLCAP];
  __shared__ u32 lused[SLCAP];
  __shared__ u32 cl[256];
  __shared__ double cw[256];
  __shared__ u64 s_head, s_tail;
  __shared__ u32 s_nused, s_best, s_v;
  __shared__ double s_wcur;
  const u32 tid = threadIdx.x;
  const Csr& g = a.g;
  const u32 n = g.n;
  const double tt = *g.twice_m;
  for (u32 q = tid; q < SLCAP; q += 256) {
    lk[q] = kNone;
    lv[q] = 0.0;
  }
  if (tid == 0) {
    s_head = 0;
    s_tail = a.qlen;
  }
  __syncthreads();
  unsigned long long moves = 0;
  while (true) {
    if (tid == 0) {
      s_v = kNone;
      if (s_head != s_tail) {
        s_v = a.queue[s_head % n];
        ++s_head;
        a.inq[s_v] = 0;
      }
      s_nused = 0;
    }
    __syncthreads();
    const u32 v = s_v;
    if (v == kNone) break;
    const u32 cur = a.f[v];
    const double kv = g.k[v];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const bool in_lds = deg <= SLCAP / 2;
    u32* HK = in_lds ? lk : a.gk;
    double* HV = in_lds ? lv : a.gv;
    u32* used = in_lds ? lused : a.gused;
    const u32 cap = in_lds ? SLCAP : a.gcap;
    for (u32 base = 0; base < deg; base += 256) {
      const u32 j = base + tid;
      if (j < deg) {
        const u32 u = g.col[off + j];
        cl[tid] = (u == v) ? kNone : a.f[u];
        cw[tid] = g.w[off + j];
      }
      __syncthreads();
      if (tid == 0) {
        const u32 m = min(256u, deg - base);
        u32 nu = s_nused;
        for (u32 q = 0; q < m; ++q) {
          if (cl[q] == kNone) continue;
          const u32 h = seq_insert(HK, cap, cl[q], used, nu);
          HV[h] += cw[q];
        }
        s_nused = nu;
      }
      __syncthreads();
    }
    if (tid == 0) s_wcur = table_lookup(HK, HV, cap, cur);
    __syncthreads();
    const double wcur = s_wcur, sc = a.cdeg[cur];
    const u32 nused = s_nused;
    double bg = 0.0;
    u32 best = kNone;
    for (u32 q = tid; q < nused; q += 256) {
      const u32 h = used[q], key = HK[h];
      if (key == cur) continue;
      const double gn = move_gain(HV[h], wcur, tt, a.gamma, kv, sc, a.cdeg[key]);
      if (gn > 0.0) better(bg, best, gn, key);
    }
    block_argmax(bg, best);
    if (tid == 0) {
      s_best = kNone;
      if (bg > 0.0 && best != kNone) {
        a.f[v] = best;
        a.cdeg[cur] -= kv;
        a.cdeg[best] += kv;
        s_best = best;
        ++moves;
      }
    }
    for (u32 q = tid; q < nused; q += 256) {
      const u32 h = used[q];
      HK[h] = kNone;
      HV[h] = 0.0;
    }
    __threadfence_block();
    __syncthreads();
    const u32 bst = s_best;
    if (bst != kNone) {
      for (u32 base = 0; base < deg; base += 256) {
        const u32 j = base + tid;
        u32 u = 0;
        bool want = false;
        if (j < deg) {
          u = g.col[off + j];
          want = u != v && a.f[u] != bst && !a.inq[u];
        }
        u32 cnt = 0;
        const u32 r = block_rank(want, &cnt);
        if (want) {
          a.queue[(s_tail + r) % n] = u;
          a.inq[u] = 1;
        }
        __threadfence_block();
        __syncthreads();
        if (tid == 0) s_tail += cnt;
        __syncthreads();
      }
    }
  }
  if (tid == 0) atomicAdd(a.moves, moves);
}

// wc[i] = weight of i's row entries inside its community (a self-loop twice),
// added to ws[sub[i]]: w(S, C) of subcommunity_delta_q_to_empty (:729-739)
__global__ void k_static_ws(Csr g, const u32* f, const u32* sub, double* wc, double* ws) {
  for (u64 v = gtid(); v < g.n; v += gstride()) {
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v], C = f[v];
    double acc = 0.0;
    for (u32 j = 0; j < deg; ++j) {
      const u32 u = g.col[off + j];
      if (f[u] == C) acc += (u == v) ? 2.0 * g.w[off + j] : g.w[off + j];
    }
    wc[v] = acc;
    if (acc != 0.0) atomicAdd(&ws[sub[v]], acc);
  }
}

struct StaticRefineArgs {
  Csr g;
  const u32* f;
  u32* sub;
  u32* size;
  double* dS;   // subcommunity degrees
  double* ws;   // w(S, C) per subcommunity
  const double* wc;
  const double* dC;  // community degrees
  const u32* order;  // the split-off vertices, degree ascending then id
  u32 nord;
  double gamma;
  u32* gk;
  double* gv;
  u32* gused;
  u32 gcap;
  u32* ckey;  // candidate lists longer than the LDS ones
  double* cval;
  u32 ccap;
};

// bitonic sort of (key, val) ascending by key over N = 2^k slots, in one block
__device__ void block_sort_pairs(u32* K, double* V, u32 N) {
  for (u32 kk = 2; kk <= N; kk <<= 1) {
    for (u32 j = kk >> 1; j > 0; j >>= 1) {
      for (u32 i = threadIdx.x; i < N; i += blockDim.x) {
        const u32 x = i ^ j;
        if (x > i) {
          const bool up = (i & kk) == 0;
          if ((K[i] > K[x]) == up) {
            const u32 tk = K[i];
            K[i] = K[x];
            K[x] = tk;
            const double tv = V[i];
            V[i] = V[x];
            V[x] = tv;
          }
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  }
}

// refine_singleton_merge's merge loop (hit_leiden.rs:617-706) in one
// workgroup: for each split-off vertex still a singleton, the weights to the
// subcommunities of its same-community neighbours (row order), candidates in
// ascending id kept by the T-filter (removal gain <= 1e-12, a zero-degree S
// kept), the argmax with the EPS rules (gain, then weight, then id), a merge
// when the best gain exceeds 1e-12.
constexpr u32 kStaticCand = 2048;
__global__ __launch_bounds__(256) void k_static_refine(StaticRefineArgs a) {
  __shared__ u32 lk[SLCAP];
  __shared__ double lv[SLCAP];
  __shared__ u32 lused[SLCAP];
  __shared__ u32 cl[256];
  __shared__ double cw[256];
  __shared__ u32 sk[kStaticCand];
  __shared__ double sv[kStaticCand];
  __shared__ u32 s_nused, s_go, s_nc;
  const u32 tid = threadIdx.x;
  const Csr& g = a.g;
  const double tt = *g.twice_m;
  const double EPS = 1e-12;
  for (u32 q = tid; q < SLCAP; q += 256) {
    lk[q] = kNone;
    lv[q] = 0.0;
  }
  __syncthreads();
  for (u32 idx = 0; idx < a.nord; ++idx) {
    const u32 v = a.order[idx];
    if (tid == 0) {
      s_go = a.size[a.sub[v]] == 1;
      s_nused = 0;
    }
    __syncthreads();
    const bool go = s_go;
    __syncthreads();  // every thread has read s_go before thread 0 writes the next one
    if (!go) continue;
    const u32 cur = a.sub[v], C = a.f[v];
    const double kv = g.k[v];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const bool in_lds = deg <= SLCAP / 2;
    u32* HK = in_lds ? lk : a.gk;
    double* HV = in_lds ? lv : a.gv;
    u32* used = in_lds ? lused : a.gused;
    const u32 cap = in_lds ? SLCAP : a.gcap;
    for (u32 base = 0; base < deg; base += 256) {
      const u32 j = base + tid;
      if (j < deg) {
        const u32 u = g.col[off + j];
        cl[tid] = (u == v || a.f[u] != C) ? kNone : a.sub[u];
        cw[tid] = g.w[off + j];
      }
      __syncthreads();
      if (tid == 0) {
        const u32 m = min(256u, deg - base);
        u32 nu = s_nused;
        for (u32 q = 0; q < m; ++q) {
          if (cl[q] == kNone) continue;
          const u32 h = seq_insert(HK, cap, cl[q], used, nu);
          HV[h] += cw[q];
        }
        s_nused = nu;
      }
      __syncthreads();
    }
    const u32 nused = s_nused;
    // the candidates (their id and weight), sorted ascending
    const bool small = nused <= kStaticCand;
    u32* CK = small ? sk : a.ckey;
    double* CV = small ? sv : a.cval;
    u32 N = 1;
    while (N < nused) N <<= 1;
    for (u32 q = tid; q < N; q += 256) {
      if (q < nused) {
        const u32 h = used[q];
        CK[q] = HK[h];
        CV[q] = HV[h];
      } else {
        CK[q] = kNone;
        CV[q] = 0.0;
      }
    }
    __threadfence_block();
    __syncthreads();
    if (nused > 1) block_sort_pairs(CK, CV, N);
    if (tid == 0) {
      double wcur = 0.0;
      for (u32 q = 0; q < nused; ++q)
        if (CK[q] == cur) wcur = CV[q];
      u32 best = cur;
      double bg = 0.0, bw = -1.0;
      const double dC = a.dC[C];
      for (u32 q = 0; q < nused; ++q) {
        const u32 S = CK[q];
        if (S == cur) continue;
        const double ds = a.dS[S];
        const double dq = ds == 0.0 ? 0.0 : (0.0 - a.ws[S]) / tt + a.gamma * ds * (dC - ds) / (tt * tt);
        if (!(dq <= EPS)) continue;
        const double w = CV[q];
        const double gain = (w - wcur) / tt + a.gamma * kv * (a.dS[cur] - kv - ds) / (tt * tt);
        const bool bt = gain > bg + EPS ||
                        (fabs(gain - bg) <= EPS && (w > bw + EPS || (fabs(w - bw) <= EPS && S < best)));
        if (bt) {
          best = S;
          bg = gain;
          bw = w;
        }
      }
      if (bg > EPS) {
        a.sub[v] = best;
        a.size[cur] -= 1;
        a.size[best] += 1;
        a.dS[cur] -= kv;
        a.dS[best] += kv;
        a.ws[cur] -= a.wc[v];
        a.ws[best] += a.wc[v];
      }
    }
    for (u32 q = tid; q < nused; q += 256) {
      const u32 h = used[q];
      HK[h] = kNone;
      HV[h] = 0.0;
    }
    __threadfence_block();
    __syncthreads();
  }
}
// active vertices of single_level_movement (:776-787): both ends of a record
// with a positive weight whose ends lie in different communities
__global__ void k_static_active(const Rec* d, u64 k, const u32* f, u32 n, u8* A) {
  for (u64 i = gtid(); i < k; i += gstride()) {
    const Rec r = d[i];
    if (r.w > 0.0 && r.u < n && r.v < n && f[r.u] != f[r.v]) {
      A[r.u] = 1;
      A[r.v] = 1;
    }
  }
}
// aggregate_graph's keys (:764-780): (min, max) of the remapped ends
__global__ void k_static_keys(const Rec* d, u64 k, const u32* remap, u64* keys, double* vals) {
  for (u64 i = gtid(); i < k; i += gstride()) {
    const Rec r = d[i];
    const u64 a = remap[r.u], b = remap[r.v];
    keys[i] = (min(a, b) << 32) | max(a, b);
    vals[i] = r.w;
  }
}
// one record per key run of the sorted keys, its weights summed in order
__global__ void k_static_reduce(const u64* keys, const double* vals, const u32* starts, u64 nruns, u64 k, Rec* out) {
  for (u64 r = gtid(); r < nruns; r += gstride()) {
    const u64 b = starts[r], e = r + 1 < nruns ? starts[r + 1] : k;
    double w = 0.0;
    for (u64 i = b; i < e; ++i) w += vals[i];
    out[r] = Rec{u32(keys[b] >> 32), u32(keys[b]), w};
  }
}
__global__ void k_heads_u64(const u64* keys, u64 n, u8* heads) {
  for (u64 i = gtid(); i < n; i += gstride()) heads[i] = (i == 0 || keys[i] != keys[i - 1]);
}
// the first member of each coarse vertex (its minimum vertex)
__global__ void k_first_member(const u32* remap, u32 n, u32* first) {
  for (u64 v = gtid(); v < n; v += gstride()) atomicMin(&first[remap[v]], u32(v));
}
__global__ void k_rep_comm(const u32* first, u32 cn, const u32* comm, u32* out) {
  for (u64 c = gtid(); c < cn; c += gstride()) out[c] = comm[first[c]];
}
__global__ void k_compose2(const u32* cf, const u32* remap, u32 n, u32* comm) {
  for (u64 v = gtid(); v < n; v += gstride()) comm[v] = cf[remap[v]];
}
__global__ void k_deg_keys(const u32* list, u64 c, const double* k, u64* keys) {
  for (u64 i = gtid(); i < c; i += gstride()) keys[i] = u64(__double_as_longlong(k[list[i]]));
}

}  // namespace

// ====================================================================== host side

#define GRID(n) blocks_for((n), 256), 256, 0, s

void fill_u8(u8* p, u8 v, u64 n, hipStream_t s) {
  HL_PROF(P_FILL, n, s);
  if (n) hipLaunchKernelGGL(k_fill_u8, GRID(n), p, v, n);
  HL_LAUNCH_CHECK();
}
void fill_u32(u32* p, u32 v, u64 n, hipStream_t s) {
  HL_PROF(P_FILL, n, s);
  if (n) hipLaunchKernelGGL(k_fill_u32, GRID(n), p, v, n);
  HL_LAUNCH_CHECK();
}
void fill_f64(double* p, double v, u64 n, hipStream_t s) {
  HL_PROF(P_FILL, n, s);
  if (n) hipLaunchKernelGGL(k_fill_f64, GRID(n), p, v, n);
  HL_LAUNCH_CHECK();
}
void iota_u32(u32* p, u32 base, u64 n, hipStream_t s) {
  HL_PROF(P_FILL, n, s);
  if (n) hipLaunchKernelGGL(k_iota_u32, GRID(n), p, base, n);
  HL_LAUNCH_CHECK();
}
void copy_u8(u8* d, const u8* src, u64 n, hipStream_t s) {
  HL_PROF(P_FILL, n, s);
  if (n) hipLaunchKernelGGL(k_copy_u8, GRID(n), d, src, n);
  HL_LAUNCH_CHECK();
}
void count4_u8(const u8* a0, const u8* a1, const u8* a2, const u8* a3, u64 n, DevBuf<u64>& cnt, u64 out[4],
               hipStream_t s) {
  cnt.reserve(4, s);
  HL_CHECK(hipMemsetAsync(cnt.p, 0, 4 * sizeof(u64), s));
  if (n) {
    HL_PROF(P_REDUCE, 4 * n, s);
    hipLaunchKernelGGL(k_count4, blocks_for(n / 4 + 1, 256, kReduceBlocks), 256, 0, s, a0, a1, a2, a3, n,
                       (unsigned long long*)cnt.p);
  }
  HL_LAUNCH_CHECK();
  d2h_async(out, cnt.p, 4 * sizeof(u64), s);
  stream_wait_reads(s);
}
void count4_u8_async(const u8* a0, const u8* a1, const u8* a2, const u8* a3, u64 n, u64* out, hipStream_t s,
                     bool cleared) {
  if (!cleared) hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, s, out, u64(0), u64(4));
  if (n) {
    HL_PROF(P_REDUCE, 4 * n, s);
    hipLaunchKernelGGL(k_count4, blocks_for(n / 4 + 1, 256, kReduceBlocks), 256, 0, s, a0, a1, a2, a3, n, (unsigned long long*)out);
  }
  HL_LAUNCH_CHECK();
}
u64 count_u8(const u8* a, u64 n, DevBuf<u64>& cnt, hipStream_t s) {
  u64 out[4];
  count4_u8(a, nullptr, nullptr, nullptr, n, cnt, out, s);
  return out[0];
}
void fill4_u8(u8* a0, u8* a1, u8* a2, u8* a3, u8 v, u64 n, hipStream_t s) {
  HL_PROF(P_FILL, 4 * n, s);
  if (n) hipLaunchKernelGGL(k_fill4_u8, GRID(n), a0, a1, a2, a3, v, n);
  HL_LAUNCH_CHECK();
}
void max_u32(const u32* a, u64 n, u32* out, hipStream_t s) {
  HL_PROF(P_REDUCE, 4 * n, s);
  if (n) hipLaunchKernelGGL(k_max_u32, blocks_for(n / 4 + 1, 256, kReduceBlocks), 256, 0, s, a, n, out);
  HL_LAUNCH_CHECK();
}
void max_u32_multi(const u32* const a[4], u64 n, u32* const out[4], hipStream_t s) {
  MaxArrays g{};
  u32 arrays = 0;
  for (int k = 0; k < 4; ++k) {
    g.a[k] = a[k];
    g.out[k] = out[k];
    arrays += a[k] != nullptr;
  }
  HL_PROF(P_REDUCE, 4 * n * arrays, s);
  if (n && arrays) hipLaunchKernelGGL(k_max_multi, blocks_for(n / 4 + 1, 256, kReduceBlocks), 256, 0, s, g, n);
  HL_LAUNCH_CHECK();
}

// ---- one-launch forms of select and scan for small inputs: each launch costs
// a few microseconds whatever its size, and the library forms take two or three
// launches; batches of a few thousand records run dozens of these.

// ordered select of up to kSelSmall flags by one workgroup of 1,024 threads
constexpr u32 kSelSmall = 1024 * kSelPer;
__global__ __launch_bounds__(1024) void k_sel_small(const u8* flags, u64 n, u32* out, unsigned long long* total) {
  __shared__ u32 wsum[16];
  u8 f[kSelPer];
  const u64 i0 = u64(threadIdx.x) * kSelPer;
  const u32 c = sel_load(flags, n, i0, f);
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u32 inc = c;
  for (int o = 1; o < 64; o <<= 1) {
    const u32 y = __shfl_up(inc, o, 64);
    if (lane >= u32(o)) inc += y;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  u32 wbase = 0;
  for (u32 i = 0; i < wv; ++i) wbase += wsum[i];
  u64 pos = wbase + inc - c;
#pragma unroll
  for (u32 j = 0; j < kSelPer; ++j)
    if (f[j]) out[pos++] = u32(i0 + j);
  if (threadIdx.x == 1023) *total = pos;
}

// exclusive scan of up to kScanSmall u64 by one workgroup, out[n] = the total
constexpr u32 kScanPer = 16, kScanSmall = 1024 * kScanPer;
__global__ __launch_bounds__(1024) void k_scan_small(const u64* in, u64* out, u64 n) {
  __shared__ u64 wsum[16];
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const u64 i0 = u64(threadIdx.x) * kScanPer;
  u64 v[kScanPer], sum = 0;
#pragma unroll
  for (u32 j = 0; j < kScanPer; ++j) {
    v[j] = i0 + j < n ? in[i0 + j] : 0;
    sum += v[j];
  }
  u64 inc = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const u32 lo = __shfl_up(u32(inc), o, 64), hi = __shfl_up(u32(inc >> 32), o, 64);
    if (lane >= u32(o)) inc += (u64(hi) << 32) | lo;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  u64 base = 0;
  for (u32 i = 0; i < wv; ++i) base += wsum[i];
  u64 run = base + inc - sum;
#pragma unroll
  for (u32 j = 0; j < kScanPer; ++j)
    if (i0 + j < n) {
      out[i0 + j] = run;
      run += v[j];
    }
  if (threadIdx.x == 1023) out[n] = base + inc;
}

void select_flagged_async(const u8* flags, u64 n, u32* out, Temp& tmp, u64* cnt_dev, hipStream_t s) {
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, s, cnt_dev, u64(0), u64(1));
    HL_LAUNCH_CHECK();
    return;
  }
  HL_PROF(P_SELECT, n, s);
  if (n <= kSelSmall) {
    hipLaunchKernelGGL(k_sel_small, 1, 1024, 0, s, flags, n, out, reinterpret_cast<unsigned long long*>(cnt_dev));
    HL_LAUNCH_CHECK();
    return;
  }
  const u64 nb = (n + kSelChunk - 1) / kSelChunk;
  u32* bcnt = static_cast<u32*>(tmp.get(nb * sizeof(u32), s));
  hipLaunchKernelGGL(k_sel_count, dim3(unsigned(nb)), dim3(256), 0, s, flags, n, bcnt);
  hipLaunchKernelGGL(k_sel_write, dim3(unsigned(nb)), dim3(256), 0, s, flags, n, bcnt, out,
                     reinterpret_cast<unsigned long long*>(cnt_dev));
  HL_LAUNCH_CHECK();
}

u64 select_flagged(const u8* flags, u64 n, u32* out, Temp& tmp, DevBuf<u64>& cnt, hipStream_t s) {
  if (n == 0) return 0;
  cnt.reserve(1, s);
  select_flagged_async(flags, n, out, tmp, cnt.p, s);
  u64 c = 0;
  d2h_async(&c, cnt.p, sizeof(u64), s);
  stream_wait_reads(s);
  return c;
}

void exclusive_scan_u64_async(const u64* in, u64* out, u64 n, Temp& tmp, hipStream_t s) {
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, s, out, u64(0), u64(1));
    HL_LAUNCH_CHECK();
    return;
  }
  HL_PROF(P_SCAN, 16 * n, s);
  if (n <= kScanSmall) {
    hipLaunchKernelGGL(k_scan_small, 1, 1024, 0, s, in, out, n);
    HL_LAUNCH_CHECK();
    return;
  }
  size_t bytes = 0;
  HL_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, u64(0), n, rocprim::plus<u64>(), s));
  void* t = tmp.get(bytes, s);
  HL_CHECK(rocprim::exclusive_scan(t, bytes, in, out, u64(0), n, rocprim::plus<u64>(), s));
  hipLaunchKernelGGL(k_scan_tail, 1, 1, 0, s, in, out, n);
  HL_LAUNCH_CHECK();
}

u64 exclusive_scan_u64(const u64* in, u64* out, u64 n, Temp& tmp, hipStream_t s) {
  if (n == 0) {
    u64 z = 0;
    HL_CHECK(hipMemcpyAsync(out, &z, sizeof(u64), hipMemcpyHostToDevice, s));
    stream_sync(s);
    return 0;
  }
  exclusive_scan_u64_async(in, out, n, tmp, s);
  u64 total = 0;
  d2h_async(&total, out + n, sizeof(u64), s);
  stream_wait_reads(s);
  return total;
}

// ---- one-launch stable sort of up to kSortSmall (key, value) pairs by one
// workgroup of 1,024 threads (the library sort takes three to five launches):
// each thread sorts 4 pairs in registers (odd-even transposition, swapping only
// on a strictly greater key), then log2(N / 4) merge stages through LDS, each
// thread placing 4 outputs from a merge-path search that takes ties from the
// left run, so equal keys keep their input order, as the radix sort does.
constexpr u32 kSortPer = 4, kSortSmall = 1024 * kSortPer;
template <class K, class V>
__global__ __launch_bounds__(1024) void k_sort_small(const K* kin, K* kout, const V* vin, V* vout, u32 n, u32 N) {
  // N: a power of two, 4 <= N <= kSortSmall, N >= n; entries n .. N-1 are
  // padding with the largest key, which stays behind every real entry
  __shared__ K ks[2][kSortSmall];
  __shared__ V vs[2][kSortSmall];
  const u32 i0 = threadIdx.x * kSortPer;
  const bool act = i0 < N;
  if (act) {
    K k[kSortPer];
    V v[kSortPer];
#pragma unroll
    for (u32 r = 0; r < kSortPer; ++r) {
      const u32 i = i0 + r;
      k[r] = i < n ? kin[i] : ~K(0);
      v[r] = i < n ? vin[i] : V();
    }
#pragma unroll
    for (u32 round = 0; round < kSortPer; ++round) {
#pragma unroll
      for (u32 a = round & 1; a + 1 < kSortPer; a += 2)
        if (k[a + 1] < k[a]) {
          const K tk = k[a];
          k[a] = k[a + 1];
          k[a + 1] = tk;
          const V tv = v[a];
          v[a] = v[a + 1];
          v[a + 1] = tv;
        }
    }
#pragma unroll
    for (u32 r = 0; r < kSortPer; ++r) {
      ks[0][i0 + r] = k[r];
      vs[0][i0 + r] = v[r];
    }
  }
  __syncthreads();
  u32 cur = 0;
  for (u32 w = kSortPer; w < N; w <<= 1) {
    if (act) {
      // runs A = [start, start + w) and B = [start + w, start + 2w); this
      // thread writes merged positions d .. d + 3 of the pair
      const u32 start = i0 & ~(2 * w - 1), d = i0 - start;
      const K* A = ks[cur] + start;
      const K* B = A + w;
      u32 lo = d > w ? d - w : 0, hi = d < w ? d : w;
      while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (!(B[d - 1 - mid] < A[mid]))
          lo = mid + 1;
        else
          hi = mid;
      }
      u32 i = lo, j = d - lo;
#pragma unroll
      for (u32 r = 0; r < kSortPer; ++r) {
        const bool take_a = i < w && (j >= w || !(B[j] < A[i]));
        const u32 from = take_a ? start + i : start + w + j;
        ks[cur ^ 1][i0 + r] = ks[cur][from];
        vs[cur ^ 1][i0 + r] = vs[cur][from];
        if (take_a)
          ++i;
        else
          ++j;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  if (act) {
#pragma unroll
    for (u32 r = 0; r < kSortPer; ++r) {
      const u32 i = i0 + r;
      if (i < n) {
        kout[i] = ks[cur][i];
        vout[i] = vs[cur][i];
      }
    }
  }
}
// the one-launch sort applies to full-width keys and n <= kSortSmall
template <class K>
static bool sort_small_ok(u64 n, int end_bit) {
  return n <= kSortSmall && end_bit == int(8 * sizeof(K));
}
template <class K, class V>
static void sort_small(const K* kin, K* kout, const V* vin, V* vout, u64 n, hipStream_t s) {
  u32 N = kSortPer;
  while (N < n) N <<= 1;
  hipLaunchKernelGGL((k_sort_small<K, V>), 1, 1024, 0, s, kin, kout, vin, vout, u32(n), N);
  HL_LAUNCH_CHECK();
}

void sort_pairs_u64(const u64* kin, u64* kout, const u32* vin, u32* vout, u64 n, int end_bit, Temp& tmp,
                    hipStream_t s) {
  if (n == 0) return;
  if (sort_small_ok<u64>(n, end_bit)) {
    HL_PROF(P_SORT, 24 * n, s);
    sort_small(kin, kout, vin, vout, n, s);
    return;
  }
  size_t bytes = 0;
  HL_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit), s));
  void* t = tmp.get(bytes, s);
  HL_PROF(P_SORT, 24 * n, s);
  HL_CHECK(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit), s));
}
void sort_pairs_u64_f64(const u64* kin, u64* kout, const double* vin, double* vout, u64 n, int end_bit, Temp& tmp,
                        hipStream_t s) {
  if (n == 0) return;
  if (sort_small_ok<u64>(n, end_bit)) {
    HL_PROF(P_SORT, 32 * n, s);
    sort_small(kin, kout, vin, vout, n, s);
    return;
  }
  size_t bytes = 0;
  HL_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit), s));
  void* t = tmp.get(bytes, s);
  HL_PROF(P_SORT, 32 * n, s);
  HL_CHECK(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit), s));
}
void sort_pairs_u32(const u32* kin, u32* kout, const u32* vin, u32* vout, u64 n, int end_bit, Temp& tmp,
                    hipStream_t s) {
  if (n == 0) return;
  if (sort_small_ok<u32>(n, end_bit)) {
    HL_PROF(P_SORT, 16 * n, s);
    sort_small(kin, kout, vin, vout, n, s);
    return;
  }
  size_t bytes = 0;
  HL_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit), s));
  void* t = tmp.get(bytes, s);
  HL_PROF(P_SORT, 16 * n, s);
  HL_CHECK(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit), s));
}

static int bits_for(u64 maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b)) ++b;
  return b;
}

void apply_delta_plan(const Rec* d, u64 k, const LevelCsrMut& g, ApplyScratch& a, Temp& tmp, ApplyPlan& plan,
                      hipStream_t s) {
  plan = ApplyPlan();
  if (k == 0) return;
  a.keys.reserve(k, s);
  a.keys_s.reserve(k, s);
  a.idx.reserve(k, s);
  a.idx_s.reserve(k, s);
  a.heads.reserve(k, s);
  a.runs.reserve(k, s);
  a.cnt.reserve(4, s);
  a.ucount.reserve(1, s);
  hipLaunchKernelGGL(k_canon_keys, GRID(k), d, k, a.keys.p, a.idx.p, a.ucount.p);
  HL_LAUNCH_CHECK();
  sort_pairs_u64(a.keys.p, a.keys_s.p, a.idx.p, a.idx_s.p, k, 64, tmp, s);
  hipLaunchKernelGGL(k_heads_u64, GRID(k), a.keys_s.p, k, 0, a.heads.p);
  HL_LAUNCH_CHECK();
  DevBuf<u64>& c1 = a.c1;
  c1.reserve(1, s);
  // the runs' count stays on the device: the resolve kernel reads it (grid
  // sized for k >= runs), and one read-back returns both counts
  select_flagged_async(a.heads.p, k, a.runs.p, tmp, c1.p, s);
  const u64 maxupd = 2 * k;
  a.ukey.reserve(maxupd, s);
  a.ukey_s.reserve(maxupd, s);
  a.uidx.reserve(maxupd, s);
  a.perm.reserve(maxupd, s);
  a.uop.reserve(maxupd, s);
  a.uw.reserve(maxupd, s);
  {
  HL_PROF(P_RESOLVE, 0, s);
  hipLaunchKernelGGL(k_resolve_runs, GRID(k), reinterpret_cast<const unsigned long long*>(c1.p), a.runs.p, k,
                     a.keys_s.p, a.idx_s.p, d, g, a.ukey.p, a.uidx.p, a.uop.p, a.uw.p, a.ucount.p);
  HL_LAUNCH_CHECK();
  }
  u32 nupd = 0;
  u64 nruns = 0;
  d2h_group({{&nupd, a.ucount.p, sizeof(u32)}, {&nruns, c1.p, sizeof(u64)}}, s);
  stream_wait_reads(s);
  plan.nruns = nruns;
  plan.nupd = nupd;
  if (g_prof && g_prof->records(P_RESOLVE)) g_prof->bytes[P_RESOLVE] += 16 * k + 32 * nruns;
  if (nupd == 0) return;
  sort_pairs_u64(a.ukey.p, a.ukey_s.p, a.uidx.p, a.perm.p, nupd, 64, tmp, s);
  a.heads.reserve(nupd, s);
  a.segs.reserve(nupd, s);
  hipLaunchKernelGGL(k_heads_u64, GRID(nupd), a.ukey_s.p, u64(nupd), 32, a.heads.p);
  HL_LAUNCH_CHECK();
  plan.nseg = select_flagged(a.heads.p, nupd, a.segs.p, tmp, c1, s);
  const u64 ns = plan.nseg;
  a.seg_row.reserve(ns, s);
  a.seg_len.reserve(ns, s);
  a.new_len.reserve(ns, s);
  a.seg_ins.reserve(ns, s);
  a.seg_del.reserve(ns, s);
  a.ins_pre.reserve(nupd, s);
  a.del_pre.reserve(nupd, s);
  a.newcap.reserve(ns, s);
  a.newoff.reserve(ns + 1, s);
  a.stlen.reserve(ns, s);
  a.stoff.reserve(ns + 1, s);
  a.chunks.reserve(ns, s);
  a.chunkbase.reserve(ns + 1, s);
  // inserts and deletes before each update, and per row, from one scan (a
  // coarse hub row can take millions of updates: no thread walks a row)
  a.ucode.reserve(nupd, s);
  a.uscan.reserve(u64(nupd) + 1, s);
  hipLaunchKernelGGL(k_upd_code, GRID(nupd), u64(nupd), a.perm.p, a.uop.p, a.ucode.p, a.cnt.p);
  HL_LAUNCH_CHECK();
  exclusive_scan_u64_async(a.ucode.p, a.uscan.p, nupd, tmp, s);
  hipLaunchKernelGGL(k_upd_pre, GRID(nupd), u64(nupd), ns, a.segs.p, a.uscan.p, a.ins_pre.p, a.del_pre.p);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_row_plan, GRID(ns), ns, a.segs.p, u64(nupd), a.ukey_s.p, a.uscan.p, g.row_len, g.row_cap,
                     a.seg_row.p, a.seg_len.p, a.new_len.p, a.seg_ins.p, a.seg_del.p, a.newcap.p, a.stlen.p,
                     a.chunks.p, a.cnt.p);
  HL_LAUNCH_CHECK();
  exclusive_scan_u64_async(a.newcap.p, a.newoff.p, ns, tmp, s);
  exclusive_scan_u64_async(a.stlen.p, a.stoff.p, ns, tmp, s);
  exclusive_scan_u64_async(a.chunks.p, a.chunkbase.p, ns, tmp, s);
  u64 h[6] = {0, 0, 0, 0, 0, 0};
  d2h_group({{h, a.cnt.p, 3 * sizeof(u64)},
             {h + 3, a.newoff.p + ns, sizeof(u64)},
             {h + 4, a.chunkbase.p + ns, sizeof(u64)},
             {h + 5, a.stoff.p + ns, sizeof(u64)}},
            s);
  stream_wait_reads(s);
  plan.old_entries = h[0];
  plan.max_row = h[1];
  plan.new_entries = h[2];
  plan.need = h[3];
  plan.total_chunks = h[4];
  plan.staged = h[5];
}

void apply_delta_commit(const LevelCsrMut& g, u64 pool_base, ApplyScratch& a, const ApplyPlan& plan, bool det,
                        const TotalsRef& tr, hipStream_t s) {
  if (plan.nseg == 0) return;
  const u64 ns = plan.nseg;
  const u32* chunk_seg = nullptr;  // block b is segment b when every row is one chunk
  if (plan.total_chunks != ns) {
    a.chunk_seg.reserve(plan.total_chunks, s);
    hipLaunchKernelGGL(k_chunk_seg, GRID(ns), ns, a.chunkbase.p, a.chunk_seg.p);
    HL_LAUNCH_CHECK();
    chunk_seg = a.chunk_seg.p;
  }
  if (plan.staged) {
    a.st_col.reserve(plan.staged, s, false, 1.25);
    a.st_w.reserve(plan.staged, s, false, 1.25);
    hipLaunchKernelGGL(k_row_stage, dim3(unsigned(plan.total_chunks)), dim3(256), 0, s, chunk_seg,
                       a.chunkbase.p, a.seg_row.p, a.stlen.p, a.stoff.p, g, a.st_col.p, a.st_w.p);
    HL_LAUNCH_CHECK();
  }
  {
  HL_PROF(P_ROW_MERGE, 12 * (plan.old_entries + plan.new_entries) + 21 * plan.nupd, s);
  hipLaunchKernelGGL(k_row_merge, dim3(unsigned(plan.total_chunks)), dim3(256), 0, s, chunk_seg, a.chunkbase.p,
                     a.segs.p,
                     a.seg_row.p, a.seg_len.p, a.seg_ins.p, a.seg_del.p, a.ukey_s.p, a.perm.p, a.uop.p, a.uw.p,
                     a.ins_pre.p, a.del_pre.p, a.newoff.p, a.newcap.p, a.stoff.p, a.st_col.p, a.st_w.p, pool_base,
                     g);
  HL_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_row_commit, GRID(ns), ns, a.seg_row.p, a.new_len.p, a.newoff.p, a.newcap.p, pool_base, g);
  HL_LAUNCH_CHECK();
  HL_PROF(P_ROW_DEGREE, 12 * plan.new_entries + 12 * ns, s);
  if (det) {
    hipLaunchKernelGGL(k_row_degree, GRID(ns), ns, a.seg_row.p, g, tr);
    HL_LAUNCH_CHECK();
  } else if (plan.max_row > kIntraHub) {
    const u32 hub_cap = u32(std::min<u64>(ns, kIntraHubList));
    const u32 pieces = u32((plan.max_row + kIntraPiece - 1) / kIntraPiece);
    a.dhubs.reserve(u64(hub_cap) + 1, s);
    a.dacc.reserve(hub_cap, s);
    fill_u32(a.dhubs.p + hub_cap, 0, 1, s);
    fill_f64(a.dacc.p, 0.0, hub_cap, s);
    hipLaunchKernelGGL(k_row_degree_wave, blocks_for(ns * 64, 256, 4096), 256, 0, s, ns, a.seg_row.p, g, tr,
                       a.dhubs.p, a.dhubs.p + hub_cap, hub_cap);
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_row_degree_hubs, 2048, 256, 0, s, g, a.dhubs.p, a.dhubs.p + hub_cap, hub_cap, pieces,
                       a.dacc.p);
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_row_degree_fin, GRID(hub_cap), g, tr, a.dhubs.p, a.dhubs.p + hub_cap, hub_cap, a.dacc.p);
    HL_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(k_row_degree_wave, blocks_for(ns * 64, 256, 4096), 256, 0, s, ns, a.seg_row.p, g, tr,
                       nullptr, nullptr, 0u);
    HL_LAUNCH_CHECK();
  }
}

u64 compact_plan(const LevelCsrMut& g, DevBuf<u64>& off_tmp, Temp& tmp, hipStream_t s, bool caps) {
  off_tmp.reserve(u64(g.n) + 1, s);
  DevBuf<u64> lens;
  lens.reserve(g.n + 1, s);
  hipLaunchKernelGGL(k_len_u64, GRID(g.n), caps ? g.row_cap : g.row_len, g.n, lens.p);
  HL_LAUNCH_CHECK();
  const u64 total = exclusive_scan_u64(lens.p, off_tmp.p, g.n, tmp, s);  // synchronises
  return total;
}

void gather_rows(const LevelCsrMut& g, const u64* offs, u32* col2, double* w2, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_rows, GRID(u64(g.n) * 64), g, offs, col2, w2);
  HL_LAUNCH_CHECK();
}

void compact_copy(const LevelCsrMut& g, const u64* offs, u32* col2, double* w2, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_rows, GRID(u64(g.n) * 64), g, offs, col2, w2);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_set_offsets, GRID(g.n), g, offs);
  HL_LAUNCH_CHECK();
  // no wait: the pools swap on the host, and every later use of either one
  // (the merges into the new pool, the next compaction into the old one, or
  // its release, which synchronises) comes after these kernels on the stream
}

u64 twice_m_scratch(u32 n) {
  u64 nb = (u64(n) + kTreeBlock - 1) / kTreeBlock;
  return 2 * nb + 2;
}

void twice_m_tree(const double* k, u32 n, double* partials, double* out, hipStream_t s) {
  HL_PROF(P_TWICE_M, 8ull * n, s);
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_f64, 1, 64, 0, s, out, 0.0, u64(1));
    HL_LAUNCH_CHECK();
    return;
  }
  const u64 nb1 = (u64(n) + kTreeBlock - 1) / kTreeBlock;
  double* bufs[2] = {partials, partials + nb1 + 1};
  const double* in = k;
  u64 cnt = n;
  int which = 0;
  while (true) {
    u64 nb = (cnt + kTreeBlock - 1) / kTreeBlock;
    double* o = (nb == 1) ? out : bufs[which];
    hipLaunchKernelGGL(k_tree_pass, dim3(unsigned(nb)), dim3(256), 0, s, in, cnt, o);
    HL_LAUNCH_CHECK();
    if (nb == 1) break;
    in = o;
    cnt = nb;
    which ^= 1;
  }
}

void grow_init(u32 n_old, u32 n_new, u32 bf, u32 bg, u32 bs, u32* f, u32* g, u32* s_pre, u32* s_cur, u64* row_off,
               u32* row_len, u32* row_cap, double* k, hipStream_t s) {
  u64 c = n_new - n_old;
  if (c) hipLaunchKernelGGL(k_grow_init, GRID(c), n_old, n_new, bf, bg, bs, f, g, s_pre, s_cur, row_off, row_len,
                            row_cap, k);
  HL_LAUNCH_CHECK();
}

void seed(const Rec* d, u64 k, const u32* f, const u32* sm, u8* A, u8* K, hipStream_t s) {
  HL_PROF(P_SEED, 24 * k, s);
  if (k) hipLaunchKernelGGL(k_seed, GRID(k), d, k, f, sm, A, K);
  HL_LAUNCH_CHECK();
}

void label_totals_atomic(const u32* lab, const double* k, u32 n, double* tot, hipStream_t s) {
  HL_PROF(P_LABEL_TOTALS, 20ull * n, s);
  if (n) hipLaunchKernelGGL(k_label_totals_atomic, GRID(n), lab, k, n, tot);
  HL_LAUNCH_CHECK();
}

void label_totals_canonical(const u32* lab, const double* k, u32 n, double* tot, u64 tot_n, DevBuf<u32>& ks,
                            DevBuf<u32>& vs, DevBuf<u32>& vidx, DevBuf<u8>& heads, DevBuf<u32>& starts,
                            DevBuf<u64>& cnt, Temp& tmp, hipStream_t s) {
  fill_f64(tot, 0.0, tot_n, s);
  if (n == 0) return;
  ks.reserve(n, s);
  vs.reserve(n, s);
  vidx.reserve(n, s);
  heads.reserve(n, s);
  starts.reserve(n, s);
  iota_u32(vidx.p, 0, n, s);
  sort_pairs_u32(lab, ks.p, vidx.p, vs.p, n, 32, tmp, s);
  hipLaunchKernelGGL(k_heads_u32, GRID(n), ks.p, u64(n), heads.p);
  HL_LAUNCH_CHECK();
  u64 nseg = select_flagged(heads.p, n, starts.p, tmp, cnt, s);
  hipLaunchKernelGGL(k_segment_sums, GRID(nseg), nseg, starts.p, u64(n), ks.p, vs.p, k, tot);
  HL_LAUNCH_CHECK();
}

int det_move(const Csr& g, u32* f, const u32* sm, double* tot, u64 tot_cap, u64* tot_len_dev, const u8* A,
             bool fresh, u8* B, u8* K, double gamma, DetMoveScratch& ds, u32 max_deg, hipStream_t s) {
  const u64 n = g.n;
  const u64 nw0 = (n + 31) / 32, nw1 = (nw0 + 31) / 32, nw2 = (nw1 + 31) / 32;
  ds.bm0.reserve(nw0 + 1, s);
  ds.bm1.reserve(nw1 + 1, s);
  ds.bm2.reserve(nw2 + 1, s);
  ds.status.reserve(1, s);
  ds.pops.reserve(1, s);
  if (fresh) {
    hipLaunchKernelGGL(k_bm_level0, GRID(nw0), A, u32(n), ds.bm0.p, nw0);
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bm_summary, GRID(nw1), ds.bm0.p, nw0, ds.bm1.p, nw1);
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bm_summary, GRID(nw2), ds.bm1.p, nw1, ds.bm2.p, nw2);
    HL_LAUNCH_CHECK();
    HL_CHECK(hipMemsetAsync(ds.pops.p, 0, sizeof(u64), s));
  }
  u32 gcap = 64;
  while (gcap < 2ull * max_deg + 2) gcap <<= 1;
  ds.gkeys.reserve(2 * u64(gcap), s);
  ds.gvals.reserve(gcap, s);
  if (max_deg > DET_LCAP / 2) {
    fill_u32(ds.gkeys.p, kNone, gcap, s);
    fill_f64(ds.gvals.p, 0.0, gcap, s);
  }
  fill_u32(ds.status.p, 0, 1, s);
  DetMoveArgs a{g, f, sm, tot, tot_cap, tot_len_dev, ds.bm0.p, ds.bm1.p, ds.bm2.p, u32(nw2), B, K, gamma,
                ds.gkeys.p, ds.gvals.p, ds.gkeys.p + gcap, gcap, ds.status.p, ds.pops.p};
  HL_PROF(P_DET_MOVE, 0, s);
  hipLaunchKernelGGL(k_det_move, 1, 256, 0, s, a);
  HL_LAUNCH_CHECK();
  u32 st = 0;
  d2h_async(&st, ds.status.p, sizeof(u32), s);
  stream_wait_reads(s);
  return int(st);
}

void sub_add_vertices(const u32* sm, u32 n0, u32 n1, u32* size, hipStream_t s) {
  if (n1 > n0) hipLaunchKernelGGL(k_sub_add, GRID(n1 - n0), sm, n0, n1, size);
  HL_LAUNCH_CHECK();
}

void sub_recompute(const u32* sm, const double* k, u32 n, u32* size, double* deg, u64 cap, hipStream_t s) {
  fill_u32(size, 0, cap, s);
  fill_f64(deg, 0.0, cap, s);
  if (n) hipLaunchKernelGGL(k_sizes, GRID(n), sm, n, size);
  HL_LAUNCH_CHECK();
  label_totals_atomic(sm, k, n, deg, s);
}

u64 cc_split(const Csr& g, u32* sm, const u8* K, u8* R, u64 s_bound, bool* identity, const SubRef& sub,
             RefineScratch& rs, Temp& tmp, hipStream_t s, const u32* f, bool first_largest, u32 maxdeg) {
  const u32 n = g.n;
  *identity = true;
  if (n == 0) return 1;
  rs.aff.reserve(s_bound, s);
  rs.status.reserve(4, s);
  rs.smax.reserve(1, s);
  // status: [0] any affected, [1] not identity, [2] max label
  hipLaunchKernelGGL(k_cc_clear, GRID(s_bound), rs.aff.p, s_bound, rs.status.p);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_aff_mark, GRID(n), K, sm, n, rs.aff.p, rs.status.p);
  HL_LAUNCH_CHECK();
  max_u32(sm, n, rs.status.p + 2, s);
  // the members of affected subcommunities are listed before the read-back,
  // so one synchronisation returns the flags, the label bound and their count
  rs.incc.reserve(n, s);
  rs.vcc.reserve(n, s);
  rs.cnt.reserve(1, s);
  hipLaunchKernelGGL(k_incc, GRID(n), sm, n, rs.aff.p, rs.incc.p);
  HL_LAUNCH_CHECK();
  select_flagged_async(rs.incc.p, n, rs.vcc.p, tmp, rs.cnt.p, s);
  u32 hs[3] = {0, 0, 0};
  u64 m = 0;
  d2h_group({{hs, rs.status.p, 3 * sizeof(u32)}, {&m, rs.cnt.p, sizeof(u64)}}, s);
  stream_wait_reads(s);
  *identity = hs[1] == 0;
  const u32 next_id = hs[2] + 1;
  if (!hs[0]) return next_id;
  if (m == 0) return next_id;
  rs.parent.reserve(n, s);
  rs.cnt32.reserve(n, s);
  rs.best64.reserve(next_id, s);
  rs.keepflag.reserve(m, s);
  rs.roots.reserve(m, s);
  rs.roots_s.reserve(m, s);
  rs.rkeys.reserve(m, s);
  rs.rkeys_s.reserve(m, s);
  rs.newid.reserve(n, s);
  HL_PROF(P_CC_SPLIT, 16 * m, s);
  {
    const u32 hub_cap = u32(std::min<u64>(m, kIntraHubList));
    const u32 pieces = u32(std::max<u64>(1, (u64(maxdeg) + kIntraPiece - 1) / kIntraPiece));
    const bool hub = maxdeg > kIntraHub;
    rs.hubs.reserve(u64(hub_cap) + 1, s);
    hipLaunchKernelGGL(k_uf_init, GRID(m), rs.vcc.p, m, rs.parent.p, rs.cnt32.p, rs.hubs.p + hub_cap);
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_uf_edges, GRID(m * 64), g, rs.vcc.p, m, sm, f, rs.parent.p, hub ? rs.hubs.p : nullptr,
                       rs.hubs.p + hub_cap, hub_cap);
    HL_LAUNCH_CHECK();
    if (hub) {
      hipLaunchKernelGGL(k_uf_edges_hubs, 2048, 256, 0, s, g, sm, f, rs.parent.p, rs.hubs.p, rs.hubs.p + hub_cap,
                         hub_cap, pieces);
      HL_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k_uf_compress, GRID(m), rs.vcc.p, m, rs.parent.p, rs.cnt32.p);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_zero_best, GRID(m), rs.vcc.p, m, sm, rs.best64.p);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_best_comp, GRID(m), rs.vcc.p, m, rs.parent.p, rs.cnt32.p, sm, rs.best64.p, first_largest);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_split_roots, GRID(m), rs.vcc.p, m, rs.parent.p, sm, rs.best64.p, rs.keepflag.p,
                     first_largest);
  HL_LAUNCH_CHECK();
  DevBuf<u32>& sel = rs.sel;
  sel.reserve(m, s);
  const u64 c = select_flagged(rs.keepflag.p, m, sel.p, tmp, rs.cnt, s);
  if (c) {
    DevBuf<u32>& vals = rs.vals;
    vals.reserve(c, s);
    hipLaunchKernelGGL(k_root_keys, GRID(c), sel.p, c, rs.vcc.p, sm, rs.rkeys.p, vals.p);
    HL_LAUNCH_CHECK();
    sort_pairs_u64(rs.rkeys.p, rs.rkeys_s.p, vals.p, rs.roots_s.p, c, 64, tmp, s);
    hipLaunchKernelGGL(k_newids, GRID(c), rs.roots_s.p, c, next_id, rs.newid.p);
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_split_assign, GRID(m), rs.vcc.p, m, rs.parent.p, rs.best64.p, rs.newid.p, sm, R, g.k, sub,
                       first_largest);
    HL_LAUNCH_CHECK();
    *identity = false;  // split ids exceed every vertex id of an identity map
  }
  return next_id + c;  // split ids are next_id .. next_id + c - 1; the kept parts keep theirs
}

void refine_parallel(const Csr& g, const u32* f, u32* sm, const u8* R, double gamma, u64 comm_space, u64 sn,
                     RefineScratch& rs, ProposeScratch& ps, MoveScratch& mvs, u32 maxdeg, const Thresholds& th,
                     Temp& tmp, const double* dpar_in, const SubRef& sub, hipStream_t s, bool keep_targets) {
  const u32 n = g.n;
  if (n == 0) return;
  // subcommunity sizes and degrees (hit_leiden.rs:1341-1357): kept by the caller, or recomputed
  u32* sizes = sub.size;
  double* dS = sub.deg;
  if (!sizes) {
    rs.sizes.reserve(sn, s);
    rs.dS.reserve(sn, s);
    sub_recompute(sm, g.k, n, rs.sizes.p, rs.dS.p, sn, s);
    sizes = rs.sizes.p;
    dS = rs.dS.p;
  }
  rs.single.reserve(n, s);
  hipLaunchKernelGGL(k_single, GRID(n), R, n, sm, sizes, rs.single.p);
  HL_LAUNCH_CHECK();
  DevBuf<u32>& sl = rs.sl;
  sl.reserve(n, s);
  const u64 ns = select_flagged(rs.single.p, n, sl.p, tmp, rs.cnt, s);
  if (ns == 0) return;
  // blocked set (compute_blocked_subcommunities)
  rs.is_cand.reserve(sn, s);
  rs.blocked.reserve(sn, s);
  rs.parentIdx.reserve(sn, s);
  rs.wS.reserve(sn, s);
  // hub-row lists of the candidate marks (hub_cap1 entries, then its count)
  // and of k_ws (hub_cap2), in one buffer reserved for both
  const u32 hub_cap1 = u32(std::min<u64>(ns, kIntraHubList));
  const u32 hub_cap2 = std::min<u32>(n, kIntraHubList);
  rs.hubs.reserve(u64(std::max(hub_cap1, hub_cap2)) + 1, s);
  hipLaunchKernelGGL(k_refine_clear, GRID(sn), rs.is_cand.p, rs.parentIdx.p, rs.wS.p, sn, rs.hubs.p + hub_cap1);
  HL_LAUNCH_CHECK();
  HL_PROF(P_REFINE, 16ull * n, s);
  {
    const u32 hub_cap = hub_cap1;
    const u32 pieces = u32(std::max<u64>(1, (u64(maxdeg) + kIntraPiece - 1) / kIntraPiece));
    const bool hub = maxdeg > kIntraHub;
    hipLaunchKernelGGL(k_cand_mark, GRID(ns * 64), g, sl.p, ns, f, sm, rs.is_cand.p, hub ? rs.hubs.p : nullptr,
                       rs.hubs.p + hub_cap, hub_cap);
    HL_LAUNCH_CHECK();
    if (hub) {
      hipLaunchKernelGGL(k_cand_mark_hubs, 2048, 256, 0, s, g, f, sm, rs.is_cand.p, rs.hubs.p, rs.hubs.p + hub_cap,
                         hub_cap, pieces);
      HL_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k_parent_idx, GRID(n), sm, n, rs.is_cand.p, rs.parentIdx.p, rs.hubs.p + hub_cap2);
  HL_LAUNCH_CHECK();
  {
    const u32 hub_cap = hub_cap2;
    const u32 pieces = u32(std::max<u64>(1, (u64(maxdeg) + kIntraPiece - 1) / kIntraPiece));
    const u64 hub_max = maxdeg > kIntraHub ? u64(pieces) * kIntraPiece : 0;
    hipLaunchKernelGGL(k_ws, blocks_for(n, 256, 4096), 256, 0, s, g, n, f, sm, rs.is_cand.p, rs.parentIdx.p,
                       rs.wS.p, rs.hubs.p, rs.hubs.p + hub_cap, hub_cap, hub_max);
    HL_LAUNCH_CHECK();
    if (hub_max) {
      hipLaunchKernelGGL(k_ws_hubs, 2048, 256, 0, s, g, f, sm, rs.parentIdx.p, rs.hubs.p, rs.hubs.p + hub_cap,
                         hub_cap, pieces, rs.wS.p);
      HL_LAUNCH_CHECK();
    }
  }
  const double* dpar = dpar_in;  // the maintained community totals, when the caller has them
  if (!dpar) {
    rs.dpar.reserve(comm_space, s);
    fill_f64(rs.dpar.p, 0.0, comm_space, s);
    label_totals_atomic(f, g.k, n, rs.dpar.p, s);
    dpar = rs.dpar.p;
  }
  hipLaunchKernelGGL(k_blocked, GRID(sn), sn, rs.is_cand.p, dS, rs.wS.p, rs.parentIdx.p, f, dpar,
                     g.twice_m, g.gid, gamma, rs.blocked.p);
  HL_LAUNCH_CHECK();
  // proposals against the snapshot, then the connected-targets rule
  (void)ps;
  (void)tmp;
  refine_proposals(g, f, sm, dS, rs.blocked.p, gamma, sl.p, ns, sn, mvs, th, maxdeg, sub, s, keep_targets);
}

void refine_det(const Csr& g, const u32* f, u32* sm, const u8* R, double gamma, u64 comm_space, u64 sn,
                RefineScratch& rs, DevBuf<u32>& ks, DevBuf<u32>& vs, DevBuf<u32>& vidx, DevBuf<u8>& heads,
                DevBuf<u32>& starts, DevBuf<u64>& cnt, u32 max_deg, Temp& tmp, hipStream_t s) {
  const u32 n = g.n;
  if (n == 0) return;
  rs.rlist.reserve(n, s);
  const u64 nr = select_flagged(R, n, rs.rlist.p, tmp, rs.cnt, s);
  if (nr == 0) return;
  rs.sizes.reserve(sn, s);
  rs.dS.reserve(sn, s);
  rs.dpar.reserve(comm_space, s);
  rs.mhead.reserve(sn, s);
  rs.mnext.reserve(n, s);
  fill_u32(rs.sizes.p, 0, sn, s);
  hipLaunchKernelGGL(k_sizes, GRID(n), sm, n, rs.sizes.p);
  HL_LAUNCH_CHECK();
  // canonical subcommunity degrees + sorted member lists from one stable sort
  label_totals_canonical(sm, g.k, n, rs.dS.p, sn, ks, vs, vidx, heads, starts, cnt, tmp, s);
  {
    // starts/ks/vs now hold the (s, v) segments
    u64 nseg = 0;
    d2h_async(&nseg, cnt.p, sizeof(u64), s);
    stream_wait_reads(s);
    fill_u32(rs.mhead.p, kNone, sn, s);
    hipLaunchKernelGGL(k_member_lists, GRID(nseg), nseg, starts.p, u64(n), ks.p, vs.p, rs.mhead.p, rs.mnext.p);
    HL_LAUNCH_CHECK();
  }
  label_totals_canonical(f, g.k, n, rs.dpar.p, comm_space, ks, vs, vidx, heads, starts, cnt, tmp, s);
  // R by (k asc, id asc)
  rs.kkeys.reserve(nr, s);
  rs.kkeys_s.reserve(nr, s);
  rs.rsorted.reserve(nr, s);
  hipLaunchKernelGGL(k_degree_keys, GRID(nr), rs.rlist.p, nr, g.k, rs.kkeys.p);
  HL_LAUNCH_CHECK();
  sort_pairs_u64(rs.kkeys.p, rs.kkeys_s.p, rs.rlist.p, rs.rsorted.p, nr, 64, tmp, s);
  u32 gcap = 64;
  while (gcap < 2ull * max_deg + 2) gcap <<= 1;
  rs.gk.reserve(2 * u64(gcap), s);
  rs.gv.reserve(gcap, s);
  if (max_deg > DET_LCAP / 2) {
    fill_u32(rs.gk.p, kNone, gcap, s);
    fill_f64(rs.gv.p, 0.0, gcap, s);
  }
  DetRefArgs a{g, f, sm, rs.rsorted.p, nr, rs.sizes.p, rs.dS.p, rs.dpar.p, rs.mhead.p, rs.mnext.p, gamma,
               rs.gk.p, rs.gv.p, rs.gk.p + gcap, gcap};
  HL_PROF(P_DET_REFINE, 0, s);
  hipLaunchKernelGGL(k_det_refine, 1, 256, 0, s, a);
  HL_LAUNCH_CHECK();
  stream_sync(s);
}

u64 aggregate(const Csr& g, const Rec* dG, u64 k, u32* s_pre, const u32* s_cur, const u8* R, DevBuf<Rec>& out,
              u32* next_n, AggScratch& as, Temp& tmp, hipStream_t s) {
  const u32 n = g.n;
  as.rlist.reserve(n + 1, s);
  DevBuf<u64>& c1 = as.c1;
  // next node count from s_cur and the old s_pre (read back with the next synchronisation)
  DevBuf<u32>& mx = as.mx;
  mx.reserve(1, s);
  fill_u32(mx.p, 0, 1, s);
  {
    const u32* const arr[4] = {s_cur, s_pre, nullptr, nullptr};
    u32* const outs[4] = {mx.p, mx.p, nullptr, nullptr};
    max_u32_multi(arr, n, outs, s);
  }
  u32 hm = 0;
  d2h_async(&hm, mx.p, sizeof(u32), s);
  const u64 nr = n ? select_flagged(R, n, as.rlist.p, tmp, c1, s) : 0;
  if (!n) stream_wait_reads(s);
  *next_n = hm + 1;
  as.cnt64.reserve(nr + 1, s);
  as.off64.reserve(nr + 1, s);
  u64 nemit = 0;
  if (nr) {
    HL_PROF(P_AGG_EMIT, 8 * nr, s);
    hipLaunchKernelGGL(k_agg_count, GRID(nr * 64), g, as.rlist.p, nr, R, as.cnt64.p);
    HL_LAUNCH_CHECK();
    nemit = exclusive_scan_u64(as.cnt64.p, as.off64.p, nr, tmp, s);
  }
  const u64 total = k + nemit;
  as.H.reserve(total + 1, s);
  if (k) hipLaunchKernelGGL(k_lift, GRID(k), dG, k, s_pre, as.H.p);
  HL_LAUNCH_CHECK();
  if (nr) {
    HL_PROF(P_AGG_EMIT, 16 * nemit + 8 * nr, s);
    hipLaunchKernelGGL(k_agg_emit, GRID(nr * 64), g, as.rlist.p, nr, R, s_pre, s_cur, as.off64.p, k, as.H.p);
  }
  HL_LAUNCH_CHECK();
  if (nr) hipLaunchKernelGGL(k_spre_update, GRID(nr), as.rlist.p, nr, s_cur, s_pre);
  HL_LAUNCH_CHECK();
  if (total == 0) return 0;
  as.keys.reserve(total, s);
  as.keys_s.reserve(total, s);
  as.idx.reserve(total, s);
  as.idx_s.reserve(total, s);
  as.heads.reserve(total, s);
  as.runs.reserve(total, s);
  hipLaunchKernelGGL(k_hkeys, GRID(total), as.H.p, total, as.keys.p, as.idx.p);
  HL_LAUNCH_CHECK();
  sort_pairs_u64(as.keys.p, as.keys_s.p, as.idx.p, as.idx_s.p, total, 64, tmp, s);
  hipLaunchKernelGGL(k_heads_u64, GRID(total), as.keys_s.p, total, 0, as.heads.p);
  HL_LAUNCH_CHECK();
  // the runs' count stays on the device (sizes use total >= runs), so the
  // survivors' count is the only read-back
  c1.reserve(1, s);
  select_flagged_async(as.heads.p, total, as.runs.p, tmp, c1.p, s);
  DevBuf<Rec>& comp = as.comp;
  comp.reserve(total, s);
  as.keep.reserve(total, s);
  {
    HL_PROF(P_COMPRESS, 33 * total, s);
    hipLaunchKernelGGL(k_compress, GRID(total), reinterpret_cast<const unsigned long long*>(c1.p), as.runs.p, total,
                       as.keys_s.p, as.idx_s.p, as.H.p, comp.p, as.keep.p);
    HL_LAUNCH_CHECK();
  }
  as.keep_idx.reserve(total, s);
  const u64 kept = select_flagged(as.keep.p, total, as.keep_idx.p, tmp, as.c2, s);
  out.reserve(kept + 1, s);
  if (kept) hipLaunchKernelGGL(k_gather_rec, GRID(kept), as.keep_idx.p, kept, comp.p, out.p);
  HL_LAUNCH_CHECK();
  return kept;
}

void def_pull(const u8* B, u32* fp, const u32* fq, const u32* sp, u32 n, double* tot, const double* k,
              hipStream_t s) {
  HL_PROF(P_DEF_UPDATE, 12ull * n, s);
  if (n) hipLaunchKernelGGL(k_def_pull, GRID(n), B, fp, fq, sp, n, tot, k);
  HL_LAUNCH_CHECK();
}
void def_push(const u8* Bp, u8* Bq, const u32* sq, u32 nq, u64 np, hipStream_t s) {
  HL_PROF(P_DEF_UPDATE, 5ull * nq, s);
  if (nq) hipLaunchKernelGGL(k_def_push, GRID(nq), Bp, Bq, sq, nq, np);
  HL_LAUNCH_CHECK();
}

void modularity_launch(const Csr& g, const u32* lab, u64 L, u32 maxdeg, double* sig, double* acc, u32* hubs,
                       u32 hub_cap, hipStream_t s) {
  const u32 n = g.n;
  fill_f64(sig, 0.0, L, s);
  fill_f64(acc, 0.0, 2, s);
  fill_u32(hubs + hub_cap, 0, 1, s);
  label_totals_atomic(lab, g.k, n, sig, s);
  HL_PROF(P_MODULARITY, 0, s);
  const u32 pieces = u32(std::max<u64>(1, (u64(maxdeg) + kIntraPiece - 1) / kIntraPiece));
  const u64 hub_max = maxdeg > kIntraHub ? u64(pieces) * kIntraPiece : 0;
  hipLaunchKernelGGL(k_intra, blocks_for(n, 256, 4096), 256, 0, s, g, lab, acc, hubs, hubs + hub_cap, hub_cap,
                     hub_max);
  HL_LAUNCH_CHECK();
  if (hub_max) {
    hipLaunchKernelGGL(k_intra_hubs, 1024, 256, 0, s, g, lab, hubs, hubs + hub_cap, hub_cap, pieces, acc);
    HL_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_sq_sum, blocks_for(L, 256, 4096), 256, 0, s, sig, L, g.twice_m, acc + 1);
  HL_LAUNCH_CHECK();
}

double modularity(const Csr& g, const u32* lab, u32 maxdeg, DevBuf<double>& sig, DevBuf<double>& acc,
                  DevBuf<u32>& mx, hipStream_t s) {
  const u32 n = g.n;
  double t = 0.0;
  d2h_async(&t, g.twice_m, sizeof(double), s);
  stream_wait_reads(s);
  if (t == 0.0 || n == 0) return 0.0;
  const u32 hub_cap = std::min<u32>(n, kIntraHubList);
  mx.reserve(u64(hub_cap) + 2, s);
  fill_u32(mx.p, 0, 1, s);
  max_u32(lab, n, mx.p, s);
  u32 hm = 0;
  d2h_async(&hm, mx.p, sizeof(u32), s);
  stream_wait_reads(s);
  const u64 L = u64(hm) + 1;
  sig.reserve(L, s);
  acc.reserve(2, s);
  modularity_launch(g, lab, L, maxdeg, sig.p, acc.p, mx.p + 1, hub_cap, s);
  double h[2];
  d2h_async(h, acc.p, 2 * sizeof(double), s);
  stream_wait_reads(s);
  return (h[0] - h[1]) / t;
}

// Presence of the labels in use: a byte per label set by plain stores (no
// atomics: every member of a community writes the same 1), then packed into
// a bitmap, 32 labels a word, with the words' popcounts.  The rank of label l
// is the popcount of the words before it (an exclusive scan over L / 32
// words) plus the bits below it in its word.
__global__ void k_mark_label_bytes(const u32* f, u32 n, u8* flag) {
  for (u64 v = gtid(); v < n; v += gstride()) flag[f[v]] = 1;
}
__global__ void k_pack_label_bits(const u8* flag, u64 W, u32* bits, u64* pc) {
  for (u64 i = gtid(); i < W; i += gstride()) {
    const uint4* q = reinterpret_cast<const uint4*>(flag + 32 * i);
    const uint4 a = q[0], b = q[1];
    const u32 wd[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    u32 m = 0;
#pragma unroll
    for (u32 j = 0; j < 32; ++j) m |= u32((wd[j >> 2] >> (8 * (j & 3))) & 1u) << j;
    bits[i] = m;
    pc[i] = u64(__popc(m));
  }
}
__device__ __forceinline__ u32 label_rank(const u32* bits, const u64* pre, u32 l) {
  return u32(pre[l >> 5]) + u32(__popc(bits[l >> 5] & ((1u << (l & 31)) - 1u)));
}
// the totals of the labels in use to their ranks (one thread per bitmap word)
__global__ void k_move_totals(const u32* bits, const u64* pre, u64 W, const double* tot, u64 tot_cap,
                              double* tot2) {
  for (u64 i = gtid(); i < W; i += gstride()) {
    u32 w = bits[i];
    u64 r = pre[i];
    while (w) {
      const u32 b = u32(__ffs(w) - 1);
      w &= w - 1;
      const u64 l = i * 32 + b;
      tot2[r++] = l < tot_cap ? tot[l] : 0.0;
    }
  }
}
__global__ void k_apply_ranks(u32* f, u32 n, const u32* bits, const u64* pre) {
  for (u64 v = gtid(); v < n; v += gstride()) f[v] = label_rank(bits, pre, f[v]);
}

void rank_labels(u32* f, u32 n, u64 L, const double* tot, u64 tot_cap, u64 tot_need, DevBuf<u64>& flags,
                 DevBuf<double>& tot2, Temp& tmp, hipStream_t s) {
  if (n == 0) return;
  const u64 W = L / 32 + 1;
  HL_PROF(P_RELABEL, 4ull * n + 20ull * W + 16ull * n, s);
  // flags: a byte per label (32 W bytes, as 4 W u64), the bitmap (W words, as
  // W / 2 + 1 u64), the word popcounts (W) and their prefix (W + 1)
  const u64 fw = 4 * W, bw = W / 2 + 1;
  flags.reserve(fw + bw + 2 * W + 1, s);
  u8* flag = reinterpret_cast<u8*>(flags.p);
  u32* bits = reinterpret_cast<u32*>(flags.p + fw);
  u64* pc = flags.p + fw + bw;
  u64* pre = pc + W;
  hipLaunchKernelGGL(k_fill_u64, GRID(fw), flags.p, u64(0), fw);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mark_label_bytes, GRID(n), f, n, flag);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pack_label_bits, GRID(W), flag, W, bits, pc);
  HL_LAUNCH_CHECK();
  exclusive_scan_u64_async(pc, pre, W, tmp, s);
  if (tot) {
    // the buffer rotates through the levels: a quarter more than this one needs
    // (relative to the need, so that the rotation does not compound it)
    tot2.reserve(std::max<u64>(tot_need, n), s, false, 1.25);
    hipLaunchKernelGGL(k_fill_f64, GRID(tot2.cap), tot2.p, 0.0, u64(tot2.cap));
    HL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_move_totals, GRID(W), bits, pre, W, tot, tot_cap, tot2.p);
    HL_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_apply_ranks, GRID(n), f, n, bits, pre);
  HL_LAUNCH_CHECK();
}

void gather_u32(u32* cur, const u32* map, u32 n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_gather_u32, GRID(n), cur, map, n);
  HL_LAUNCH_CHECK();
}

void canonicalize(u32* lab, u32 n, DevBuf<u32>& first, DevBuf<u8>& flag, DevBuf<u64>& pos, Temp& tmp,
                  hipStream_t s) {
  if (n == 0) return;
  DevBuf<u32> mx;
  mx.reserve(1, s);
  fill_u32(mx.p, 0, 1, s);
  max_u32(lab, n, mx.p, s);
  u32 hm = 0;
  d2h_async(&hm, mx.p, sizeof(u32), s);
  stream_wait_reads(s);
  const u64 L = u64(hm) + 1;
  first.reserve(L, s);
  fill_u32(first.p, kNone, L, s);
  hipLaunchKernelGGL(k_first_seen, GRID(n), lab, n, first.p);
  HL_LAUNCH_CHECK();
  DevBuf<u64> fl;
  fl.reserve(n + 1, s);
  pos.reserve(n + 1, s);
  hipLaunchKernelGGL(k_first_flags, GRID(n), lab, n, first.p, fl.p);
  HL_LAUNCH_CHECK();
  exclusive_scan_u64(fl.p, pos.p, n, tmp, s);
  hipLaunchKernelGGL(k_relabel, GRID(n), lab, n, first.p, pos.p);
  HL_LAUNCH_CHECK();
  (void)flag;
}


// ---- several graphs per engine ----
void gid_from_offsets(const u32* offs, u32 G, u32 n, u32* gid, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_gid_from_offsets, GRID(n), offs, G, n, gid);
  HL_LAUNCH_CHECK();
}
void graph_counts(const Rec* d, u64 k, const u32* gid, u32 G, u32* cnt, hipStream_t s) {
  fill_u32(cnt, 0, u64(G) + 1, s);
  if (k) hipLaunchKernelGGL(k_graph_counts, GRID(k), d, k, gid, cnt);
  HL_LAUNCH_CHECK();
}
void list_graph_counts(const u32* list, u64 c, const u32* gid, u32 G, u32* cnt, hipStream_t s) {
  fill_u32(cnt, 0, u64(G) + 1, s);
  if (c) hipLaunchKernelGGL(k_list_graph_counts, GRID(c), list, c, gid, cnt);
  HL_LAUNCH_CHECK();
}
void part_from_counts(const u32* cnt, u32 G, u8* part, hipStream_t s) {
  hipLaunchKernelGGL(k_part_from_counts, GRID(u64(G) + 1), cnt, G, part);
  HL_LAUNCH_CHECK();
}
void identity_graphs_to_R(const u32* sm, const u32* gid, u32 n, u32 G, const u8* part, u8* nid, u8* R,
                          hipStream_t s) {
  fill_u8(nid, 0, u64(G) + 1, s);
  if (n) {
    hipLaunchKernelGGL(k_graph_nonident, GRID(n), sm, gid, n, nid);
    hipLaunchKernelGGL(k_fill_R_graphs, GRID(n), gid, n, part, nid, R);
  }
  HL_LAUNCH_CHECK();
}
void scatter_gid(const u32* sm, const u32* gid_p, u32 n, u32 G, u32* gid_q, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_scatter_gid, GRID(n), sm, gid_p, n, G, gid_q);
  HL_LAUNCH_CHECK();
}
void modularity_graphs(const Csr& g, const u32* lab, u32 G, double* out, DevBuf<double>& sig, DevBuf<u32>& cg,
                       DevBuf<double>& acc, DevBuf<u32>& mx, hipStream_t s) {
  const u32 n = g.n;
  std::vector<double> h(2 * (u64(G) + 1), 0.0), tm(u64(G) + 1, 0.0);
  mx.reserve(1, s);
  fill_u32(mx.p, 0, 1, s);
  if (n) max_u32(lab, n, mx.p, s);
  u32 hm = 0;
  d2h_async(&hm, mx.p, sizeof(u32), s);
  stream_wait_reads(s);
  const u64 L = u64(hm) + 1;
  sig.reserve(L, s);
  cg.reserve(L, s);
  acc.reserve(2 * (u64(G) + 1), s);
  fill_f64(sig.p, 0.0, L, s);
  fill_u32(cg.p, G, L, s);
  fill_f64(acc.p, 0.0, 2 * (u64(G) + 1), s);
  if (n) {
    HL_PROF(P_MODULARITY, 0, s);
    hipLaunchKernelGGL(k_sig_graph, GRID(n), lab, g.k, g.gid, n, sig.p, cg.p);
    const u64 chunk = 64;
    hipLaunchKernelGGL(k_intra_graphs, GRID((n + chunk - 1) / chunk), g, lab, chunk, acc.p);
    hipLaunchKernelGGL(k_sq_graphs, GRID((L + chunk - 1) / chunk), sig.p, cg.p, L, chunk, acc.p + G + 1);
  }
  HL_LAUNCH_CHECK();
  HL_CHECK(hipMemcpyAsync(h.data(), acc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HL_CHECK(hipMemcpyAsync(tm.data(), g.twice_m, tm.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  stream_sync(s);
  // compute_modularity (hit_leiden.rs:121-161) per graph
  for (u32 i = 0; i < G; ++i) {
    const double t = tm[i];
    out[i] = t == 0.0 ? 0.0 : (h[i] - h[G + 1 + i] / t) / t;
  }
}


// contingency_metrics (hit_leiden_incremental.rs:577-666) on the device
void compare_partitions(const u32* a, const u32* b, u64 n, u32 la, u32 lb, double out[4], Temp& tmp,
                        hipStream_t s) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (n == 0) return;
  DevBuf<u32> mx;
  mx.reserve(2, s);
  fill_u32(mx.p, 0, 2, s);
  {
    const u32* const arr[4] = {a, b, nullptr, nullptr};
    u32* const outs[4] = {mx.p, mx.p + 1, nullptr, nullptr};
    max_u32_multi(arr, n, outs, s);
  }
  u32 hm[2];
  d2h_async(hm, mx.p, sizeof(hm), s);
  stream_wait_reads(s);
  const u64 LA = u64(hm[0]) + 1, LB = u64(hm[1]) + 1;
  DevBuf<u32> sa, sb, besta, bestb, idx, idx_s, starts, misc;
  DevBuf<u64> keys, keys_s, cnt;
  DevBuf<u8> heads;
  DevBuf<double> acc;
  DevBuf<unsigned long long> pur;
  sa.reserve(LA, s);
  sb.reserve(LB, s);
  besta.reserve(LA, s);
  bestb.reserve(LB, s);
  fill_u32(sa.p, 0, LA, s);
  fill_u32(sb.p, 0, LB, s);
  fill_u32(besta.p, 0, LA, s);
  fill_u32(bestb.p, 0, LB, s);
  keys.reserve(n, s);
  keys_s.reserve(n, s);
  idx.reserve(n, s);
  idx_s.reserve(n, s);
  heads.reserve(n, s);
  starts.reserve(n, s);
  hipLaunchKernelGGL(k_pair_keys, GRID(n), a, b, n, keys.p, sa.p, sb.p);
  HL_LAUNCH_CHECK();
  iota_u32(idx.p, 0, n, s);
  sort_pairs_u64(keys.p, keys_s.p, idx.p, idx_s.p, n, 64, tmp, s);
  hipLaunchKernelGGL(k_heads_u64, GRID(n), keys_s.p, n, 0, heads.p);
  HL_LAUNCH_CHECK();
  const u64 nruns = select_flagged(heads.p, n, starts.p, tmp, cnt, s);
  acc.reserve(3, s);
  fill_f64(acc.p, 0.0, 3, s);
  misc.reserve(1, s);
  fill_u32(misc.p, 0, 1, s);
  pur.reserve(2, s);
  HL_CHECK(hipMemsetAsync(pur.p, 0, 2 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_run_terms, GRID(nruns), keys_s.p, starts.p, nruns, n, sa.p, sb.p, besta.p, bestb.p, la, lb,
                     acc.p, misc.p);
  hipLaunchKernelGGL(k_label_terms, GRID(LA), sa.p, besta.p, LA, n, acc.p + 1, pur.p);
  hipLaunchKernelGGL(k_label_terms, GRID(LB), sb.p, bestb.p, LB, n, acc.p + 2, pur.p + 1);
  HL_LAUNCH_CHECK();
  double h[3];
  unsigned long long p[2];
  u32 inter = 0, szl[2] = {0, 0};
  d2h_async(h, acc.p, sizeof(h), s);
  d2h_async(p, pur.p, sizeof(p), s);
  d2h_async(&inter, misc.p, sizeof(u32), s);
  if (la < LA) d2h_async(&szl[0], sa.p + la, sizeof(u32), s);
  if (lb < LB) d2h_async(&szl[1], sb.p + lb, sizeof(u32), s);
  stream_sync(s);
  out[0] = (h[1] > 0.0 && h[2] > 0.0) ? h[0] / std::sqrt(h[1] * h[2]) : 0.0;
  out[1] = double(p[0]) / double(n);
  out[2] = double(p[1]) / double(n);
  const u64 uni = u64(szl[0]) + szl[1] - inter;
  out[3] = uni ? double(inter) / double(uni) : 0.0;
}


// ------------------------------------------------------------------ static path (host side)
void static_active(const Rec* d, u64 k, const u32* f, u32 n, u8* A, hipStream_t s) {
  fill_u8(A, 0, n, s);
  if (k) hipLaunchKernelGGL(k_static_active, GRID(k), d, k, f, n, A);
  else fill_u8(A, 1, n, s);
  HL_LAUNCH_CHECK();
}

void static_movement(const Csr& g, u32* f, double* cdeg, const u32* queue_host, u32 qlen, const u8* active,
                     u32 maxdeg, double gamma, StaticScratch& ss, hipStream_t s) {
  const u32 n = g.n;
  if (n == 0) return;
  ss.queue.reserve(n + 1, s);
  ss.inq.reserve(n + 1, s);
  ss.moves.reserve(1, s);
  copy_u8(ss.inq.p, active, n, s);
  if (qlen) HL_CHECK(hipMemcpyAsync(ss.queue.p, queue_host, qlen * sizeof(u32), hipMemcpyHostToDevice, s));
  u32 gcap = 1;
  while (gcap < 2 * std::max<u32>(maxdeg, 1)) gcap <<= 1;
  if (maxdeg > SLCAP / 2) {
    ss.gk.reserve(gcap, s);
    ss.gv.reserve(gcap, s);
    ss.gused.reserve(gcap, s);
    fill_u32(ss.gk.p, kNone, gcap, s);
    fill_f64(ss.gv.p, 0.0, gcap, s);
  }
  StaticMoveArgs a{g, f, cdeg, ss.queue.p, qlen, ss.inq.p, gamma, ss.gk.p, ss.gv.p, ss.gused.p, gcap, ss.moves.p};
  HL_PROF(P_DET_MOVE, 0, s);
  hipLaunchKernelGGL(k_static_move, 1, 256, 0, s, a);
  HL_LAUNCH_CHECK();
}

void static_refine_merge(const Csr& g, const u32* f, u32* sub, u32* size, double* dS, u64 sspace, const double* dC,
                         const u8* R, u32 maxdeg, double gamma, StaticScratch& ss, Temp& tmp, hipStream_t s) {
  const u32 n = g.n;
  if (n == 0) return;
  // w(S, C) per subcommunity from each vertex's weight inside its community
  ss.wc.reserve(n, s);
  ss.ws.reserve(sspace, s);
  fill_f64(ss.ws.p, 0.0, sspace, s);
  hipLaunchKernelGGL(k_static_ws, GRID(n), g, f, sub, ss.wc.p, ss.ws.p);
  HL_LAUNCH_CHECK();
  // the split-off vertices by degree ascending, then id (a stable sort of the id-ordered list)
  ss.list.reserve(n, s);
  const u64 c = select_flagged(R, n, ss.list.p, tmp, ss.cnt, s);
  if (c == 0) return;
  ss.keys.reserve(c, s);
  ss.keys2.reserve(c, s);
  ss.order.reserve(c, s);
  hipLaunchKernelGGL(k_deg_keys, GRID(c), ss.list.p, c, g.k, ss.keys.p);
  HL_LAUNCH_CHECK();
  sort_pairs_u64(ss.keys.p, ss.keys2.p, ss.list.p, ss.order.p, c, 64, tmp, s);
  u32 gcap = 1;
  while (gcap < 2 * std::max<u32>(maxdeg, 1)) gcap <<= 1;
  if (maxdeg > SLCAP / 2) {
    ss.gk.reserve(gcap, s);
    ss.gv.reserve(gcap, s);
    ss.gused.reserve(gcap, s);
    fill_u32(ss.gk.p, kNone, gcap, s);
    fill_f64(ss.gv.p, 0.0, gcap, s);
  }
  u32 ccap = 1;
  while (ccap < std::max<u32>(maxdeg, 1)) ccap <<= 1;
  if (maxdeg > kStaticCand) {
    ss.ckey.reserve(ccap, s);
    ss.cval.reserve(ccap, s);
  }
  StaticRefineArgs a{g, f, sub, size, dS, ss.ws.p, ss.wc.p, dC, ss.order.p, u32(c), gamma, ss.gk.p, ss.gv.p,
                     ss.gused.p, gcap, ss.ckey.p, ss.cval.p, ccap};
  HL_PROF(P_DET_REFINE, 0, s);
  hipLaunchKernelGGL(k_static_refine, 1, 256, 0, s, a);
  HL_LAUNCH_CHECK();
}

u64 static_aggregate(const Rec* d, u64 k, const u32* remap, Rec* out, StaticScratch& ss, Temp& tmp, hipStream_t s) {
  if (k == 0) return 0;
  ss.keys.reserve(k, s);
  ss.keys2.reserve(k, s);
  ss.vals.reserve(k, s);
  ss.vals2.reserve(k, s);
  ss.heads.reserve(k, s);
  ss.starts.reserve(k, s);
  hipLaunchKernelGGL(k_static_keys, GRID(k), d, k, remap, ss.keys.p, ss.vals.p);
  HL_LAUNCH_CHECK();
  sort_pairs_u64_f64(ss.keys.p, ss.keys2.p, ss.vals.p, ss.vals2.p, k, 64, tmp, s);
  hipLaunchKernelGGL(k_heads_u64, GRID(k), ss.keys2.p, k, ss.heads.p);
  HL_LAUNCH_CHECK();
  const u64 nruns = select_flagged(ss.heads.p, k, ss.starts.p, tmp, ss.cnt, s);
  hipLaunchKernelGGL(k_static_reduce, GRID(nruns), ss.keys2.p, ss.vals2.p, ss.starts.p, nruns, k, out);
  HL_LAUNCH_CHECK();
  return nruns;
}

void static_coarse_start(const u32* remap, u32 n, u32 cn, const u32* comm, u32* out, StaticScratch& ss,
                         hipStream_t s) {
  ss.first.reserve(cn + 1, s);
  fill_u32(ss.first.p, kNone, cn, s);
  hipLaunchKernelGGL(k_first_member, GRID(n), remap, n, ss.first.p);
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_rep_comm, GRID(cn), ss.first.p, cn, comm, out);
  HL_LAUNCH_CHECK();
}

void static_compose(const u32* cf, const u32* remap, u32 n, u32* comm, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_compose2, GRID(n), cf, remap, n, comm);
  HL_LAUNCH_CHECK();
}

}  // namespace hl
