// hl_move.hip -- throughput-mode local moving with device-side frontiers.
//
// One round (inc_movement_parallel, throughput.rs:80-202):
//   k_prop2_first  split the active list by degree (wave-aggregated appends) and
//                  evaluate the low-degree vertices in the same pass
//   k_prop2_*      per-vertex argmax over neighbour communities (parallel_frontier.rs:155-236);
//                  a positive proposal is appended to the candidate list
//   k_dec_small    one workgroup: sort candidates by (gain desc, node asc),
//                  decoupled priority rounds (throughput.rs:204-236), apply the
//                  accepted moves, mark next-active neighbours and K
//                  (throughput.rs:182-199) -- no host round trip inside a round
// Large candidate sets (the initial pass) take the multi-kernel path.
// Lists are deduplicated with a bitmap of claimed vertices, cleared per round.
#include <hip/hip_cooperative_groups.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cstring>

#include "hl_move.h"

namespace hl {
namespace {

constexpr u32 SMALL_MAX2 = 16;
constexpr u32 WAVE_HW2 = 1024;  // table slots of a wavefront-per-vertex proposal
constexpr u32 kHalfDeg = 128;   // half-wavefront-per-vertex proposals up to this degree
constexpr u32 kUndec = 0, kAcc = 1, kRej = 2;
constexpr u32 kDecLds = 1024;   // candidates whose priority rounds run on LDS slots
constexpr u32 kDecHash = 4096;  // LDS hash slots for their 2 x kDecLds communities
constexpr u32 kWaveBlocks = 768;
constexpr u32 kHalfBlocks = 1280;
constexpr u32 kTabBlocks = 2048;  // small-table proposals: up to 8 workgroups per CU
static_assert(4 + kBuckets <= 13, "bucket counts must end before the flat mark count");
// counters[kOvfSlot + i]: entries of overflow list i (vertices whose distinct
// neighbour keys did not fit the small tables of their bucket)
constexpr u32 kOvfSlot = 16, kOvfLists = 4;
// bucket counts start at counters[4]
// list index (and count slot counters[4 + kHtList]) of the vertices evaluated
// from their hub tables
constexpr u32 kHtList = kBuckets;
static_assert(4 + kHtList < 12, "the table-list count must end before the cache count");
__device__ __forceinline__ void clear_bucket_counts(u32* bucket_counts) {
  for (u32 i = 0; i <= kHtList; ++i) bucket_counts[i] = 0;
  for (u32 i = 0; i < kOvfLists; ++i) bucket_counts[kOvfSlot - 4 + i] = 0;
}
constexpr u32 kLdsDeg = 4096;       // workgroup-per-vertex with an LDS table up to this degree
constexpr u32 kLdsCap = 8192;       // its table slots (big_cap(kLdsDeg))
constexpr u32 kBigDeg = 1u << 18;   // workgroup-per-vertex with an HBM table up to this degree
constexpr u64 kMarksInBlock = 2048;  // neighbour entries marked inside the decoupling workgroup
constexpr unsigned kMarksBlocks = 1024;
constexpr u32 kMarkPer = 4;  // marks: row entries per thread and pass of k_marks_flat
constexpr u64 kBigTableBytes = 3ull << 30;  // HBM hub tables of the overflow proposals
constexpr u32 kOvfBlocks = 64;              // workgroups (and HBM tables) for the overflow list

__device__ __forceinline__ u64 gtid() { return u64(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ __forceinline__ u64 gstride() { return u64(gridDim.x) * blockDim.x; }
// integer mix (lowbias32) so that the low bits used as a table index depend on every key bit
__device__ __forceinline__ u32 hash32(u32 x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ double move_gain(double wc, double wcur, double t, double gamma, double k, double sc,
                                            double scand) {
  return (wc - wcur) / t + gamma * k * (sc - k - scand) / (t * t);
}
__device__ __forceinline__ double empty_gain(double wcur, double t, double gamma, double k, double sc) {
  return -wcur / t + gamma * k * (sc - k) / (t * t);
}
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

// Append from the lanes that want to; one atomic per wavefront.
__device__ __forceinline__ u32 wave_append(bool want, u32* cnt) {
  unsigned long long m = __ballot(want);
  if (m == 0) return kNone;
  const u32 lane = threadIdx.x & 63;
  const u32 leader = u32(__ffsll((long long)m) - 1);
  u32 base = 0;
  if (lane == leader) base = atomicAdd(cnt, u32(__popcll(m)));
  base = __shfl(base, leader, 64);
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  return want ? base + u32(__popcll(m & below)) : kNone;
}
// The same with one device atomic per workgroup pass; every thread of the
// block must call it, in step (blockDim.x <= 1024).
__device__ __forceinline__ u32 block_append(bool want, u32* cnt) {
  __shared__ u32 bw[16], bo[16];
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const unsigned long long m = __ballot(want);
  if (lane == 0) bw[wv] = u32(__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 tot = 0;
    for (u32 w = 0; w < nw; ++w) tot += bw[w];
    u32 base = tot ? atomicAdd(cnt, tot) : 0;
    for (u32 w = 0; w < nw; ++w) {
      bo[w] = base;
      base += bw[w];
    }
  }
  __syncthreads();
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const u32 p = bo[wv] + u32(__popcll(m & below));
  __syncthreads();  // bw and bo are reused by the next call
  return want ? p : kNone;
}
// load that observes stores made earlier in the same kernel by other workgroup threads
__device__ __forceinline__ u32 ld_agent(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// the store such a load observes once a barrier orders the two (an agent-scope
// atomic store needs no device-wide fence to be seen by ld_agent)
__device__ __forceinline__ void st_agent(u32* p, u32 x) {
  __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Slot of key in an LDS open-addressing table (linear probing, inserts if absent).
__device__ __forceinline__ u32 lds_insert(u32* K, u32 cap, u32 key) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    const u32 old = atomicCAS(&K[h], kNone, key);
    if (old == kNone || old == key) return h;
    h = (h + 1) & (cap - 1);
  }
}
// Add v to a list once per round: one bit per vertex in a bitmap cleared when
// the round starts (n / 8 bytes, which stay in L2, where a per-vertex stamp
// costs a line of HBM traffic per random test).  A plain read first spares the
// atomic for vertices already claimed, which hubs make common.
__device__ __forceinline__ bool marked(const u32* bits, u32 v) { return (bits[v >> 5] >> (v & 31)) & 1u; }
__device__ __forceinline__ bool claim(u32* bits, u32 v) {
  u32* w = bits + (v >> 5);
  const u32 b = 1u << (v & 31);
  if (*w & b) return false;
  return (atomicOr(w, b) & b) == 0;
}
__device__ __forceinline__ u64 ld_agent64(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// the f half of fs[v]
__device__ __forceinline__ u32* fs_f(u64* fs, u32 v) { return reinterpret_cast<u32*>(fs + v); }
__global__ void k_pack_fs(const u32* f, const u32* s, u32 n, u64* fs) {
  for (u64 v = gtid(); v < n; v += gstride()) fs[v] = (u64(s[v]) << 32) | f[v];
}
// Keep the neighbour-community cache current after v moved to community to:
// the entry for v in neighbour u's row (rows sorted by neighbour id) gets to.
__device__ __forceinline__ void ck_update(const MoveCtx& m, u32 u, u32 v, u32 to) {
  if (!m.ck || u == v) return;
  const u64 off = m.g.row_off[u];
  const u32 len = m.g.row_len[u];
  u32 lo = 0, hi = len;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (m.g.col[off + mid] < v) lo = mid + 1;
    else hi = mid;
  }
  if (lo < len && m.g.col[off + lo] == v) m.ck[off + lo] = to;
}
// Rebuild the cache for every row: one lane per row up to 64 entries, the
// whole wavefront for a longer one.
__global__ void k_key_cache(Csr g, const u32* f, u32* ck) {
  const u32 lane = threadIdx.x & 63;
  for (u64 base = gtid() - lane; base < g.n; base += gstride()) {
    const u64 v = base + lane;
    u64 off = 0;
    u32 deg = 0;
    if (v < g.n) {
      off = g.row_off[v];
      deg = g.row_len[v];
    }
    const bool wide = deg > 64;
    if (!wide)
      for (u32 j = 0; j < deg; ++j) {
        const u32 u = g.col[off + j];
        ck[off + j] = u == v ? kNone : f[u];
      }
    unsigned long long wm = __ballot(wide);
    while (wm) {
      const int src = __ffsll((long long)wm) - 1;
      wm &= wm - 1;
      const u64 bv = __shfl((unsigned long long)v, src, 64);
      const u64 boff = __shfl((unsigned long long)off, src, 64);
      const u32 bdeg = __shfl(deg, src, 64);
      for (u32 j = lane; j < bdeg; j += 64) {
        const u32 u = g.col[boff + j];
        ck[boff + j] = u == bv ? kNone : f[u];
      }
    }
  }
}
// Workgroup-table slots for a hub of degree deg (load factor at most 2/3).
__host__ __device__ inline u32 big_cap(u32 deg) {
  u32 c = 64;
  while (c < deg + (deg >> 1)) c <<= 1;
  return c;
}

// Slot of key in an LDS table of cap slots (linear probing, inserts if absent).
__device__ __forceinline__ u32 tab_insert(u32* K, u32 cap, u32 key) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    u32 old = atomicCAS(&K[h], kNone, key);
    if (old == kNone || old == key) return h;
    h = (h + 1) & (cap - 1);
  }
}
// The same for the HBM tables, with atomics only: a plain read there may be
// served by the CU's L1, which device atomics do not update.
__device__ __forceinline__ u32 hbm_insert(u32* K, u32 cap, u32 key) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    const u32 old = atomicCAS(&K[h], kNone, key);
    if (old == kNone || old == key) return h;
    h = (h + 1) & (cap - 1);
  }
}
__device__ __forceinline__ double tab_lookup(const u32* K, const double* V, u32 cap, u32 key) {
  u32 h = hash32(key) & (cap - 1);
  while (true) {
    u32 x = K[h];
    if (x == key) return V[h];
    if (x == kNone) return 0.0;
    h = (h + 1) & (cap - 1);
  }
}

// ------------------------------------------------------------------ seeding
__global__ void k_seed2(const Rec* d, u64 k, const u32* f, const u32* s, u8* A, u8* K, u32* mark,
                        u32* list, u32* cnt) {
  for (u64 base = u64(blockIdx.x) * blockDim.x; base < k; base += gstride()) {  // uniform per workgroup
    const u64 i = base + threadIdx.x;
    bool cu = false, cv = false;
    Rec r{0, 0, 0.0};
    if (i < k) {
      r = d[i];
      const bool same_f = f[r.u] == f[r.v];
      const bool act = (r.w > 0.0 && !same_f) || (r.w < 0.0 && same_f);
      const bool kk = s[r.u] == s[r.v];
      cu = act && claim(mark, r.u);
      cv = act && r.v != r.u && claim(mark, r.v);
      if (act) {
        A[r.u] = 1;
        A[r.v] = 1;
      }
      if (kk) {
        K[r.u] = 1;
        K[r.v] = 1;
      }
    }
    const u32 pu = block_append(cu, cnt);
    if (cu) list[pu] = r.u;
    const u32 pv = block_append(cv, cnt);
    if (cv) list[pv] = r.v;
  }
}
// vertices of taking-part graphs without records at this level
__global__ void k_activate_graphs(const u32* gid, u32 n, const u8* part, const u32* gcnt, u8* A, u32* mark,
                                  u32* list, u32* cnt) {
  for (u64 base = u64(blockIdx.x) * blockDim.x; base < n; base += gstride()) {
    const u64 v = base + threadIdx.x;
    bool want = false;
    if (v < n) {
      const u32 g = gid[v];
      if (part[g] && gcnt[g] == 0) {
        A[v] = 1;
        want = claim(mark, u32(v));
      }
    }
    const u32 p = block_append(want, cnt);
    if (want) list[p] = u32(v);
  }
}
__global__ void k_fill_u64_one(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}
// the movement's 32 counters and its claim bitmap, cleared
__global__ void k_move_clear(u32* counters, u32* mark, u64 words) {
  if (gtid() < 32) counters[gtid()] = 0;
  for (u64 w = gtid(); w < words; w += gstride()) mark[w] = 0;
}
__global__ void k_iota_list(u32* list, u32 n, u32* cnt) {
  for (u64 i = gtid(); i < n; i += gstride()) list[i] = u32(i);
  if (gtid() == 0) *cnt = n;
}

// One device atomic per workgroup (same-address atomics serialise); every
// thread of the block must call it.
__device__ void block_add_u64(unsigned long long* dst, unsigned long long x) {
  __shared__ unsigned long long part[16];
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) part[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (u32 i = 0; i < nw; ++i) t += part[i];
    if (t) atomicAdd(dst, t);
  }
}

// K counters at once: dst[idx[i]] += vals[i], one device atomic per counter and
// workgroup; every thread of the block must call it (blockDim <= 1024).
template <int K>
__device__ void block_add_multi(unsigned long long* dst, const u32 (&idx)[K], unsigned long long (&vals)[K]) {
  __shared__ unsigned long long part[16][K];
  const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    unsigned long long x = vals[i];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    vals[i] = x;
  }
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < K; ++i) part[wv][i] = vals[i];
  __syncthreads();
  if (threadIdx.x < u32(K)) {
    unsigned long long t = 0;
    for (u32 w = 0; w < nw; ++w) t += part[w][threadIdx.x];
    if (t) atomicAdd(&dst[idx[threadIdx.x]], t);
  }
}

// ------------------------------------------------------------------ proposals
template <class P>
__device__ __forceinline__ void finish_choice(const P& pol, u32 v, u32 cur, double wcur, double kv, double t,
                                              double& bg, u32& best) {
  if constexpr (P::kEmpty) {
    double eg = empty_gain(wcur, t, pol.gamma, kv, pol.cur_total(cur));
    if (eg > bg) {
      bg = eg;
      best = pol.empty_id(v);
    }
  }
}

// The round's first kernel: one thread per active vertex; vertices up to
// small_deg are evaluated here (row-order sums per key in registers), the
// others are appended to the bucket lists of the kernels that follow (half
// wavefront <= half_deg, wavefront <= wave_deg, LDS workgroup <= kLdsDeg, hub
// workgroup <= kBigDeg, giant).
template <class P>
__global__ __launch_bounds__(256) void k_prop2_first(P pol, Csr g, const u32* list, const u32* cnt_dev,
                                                      u32 small_deg, u32 half_deg, u32 wave_deg, u32* lists, u64 cap,
                                                      u32* counts,
                                                      CandOut out, unsigned long long* st) {
  const u32 cnt = *cnt_dev;
  const u32 t = threadIdx.x;
  unsigned long long vals[2 * kBuckets + 1] = {};  // per bucket: vertices, row entries; small-bucket candidates
  for (u64 base = u64(blockIdx.x) * 256; base < cnt; base += gstride()) {
    const u64 id = base + t;
    double bg = 0.0;
    u32 best = kNone, v = 0, b = 0xFFu, deg = 0;  // b: bucket, kHtList, or none
    if (id < cnt) {
      v = list[id];
      deg = g.row_len[v];
      b = deg <= small_deg ? 0
          : deg <= half_deg ? 1
          : deg <= wave_deg ? 2
          : deg <= kLdsDeg  ? 3
          : deg <= kBigDeg  ? 4
                            : 5;
      if constexpr (P::kTables)
        if (b >= 3 && pol.table_ok(v)) b = kHtList;  // counted by k_prop2_ht
#pragma unroll
      for (u32 q = 0; q < kBuckets; ++q)
        if (q == b) {
          vals[q] += 1;
          vals[kBuckets + q] += deg;
        }
    }
    {
      // the bucket lists 1 .. kHtList: one device atomic per list and
      // workgroup pass (same-address atomics serialise, so one per wavefront
      // costs milliseconds when millions of vertices are active)
      constexpr u32 NL = kHtList;  // lists 1 .. NL
      __shared__ u32 wcnt[NL][4], woff[NL][4];
      const u32 lane = t & 63, wv = t >> 6;
      const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
      unsigned long long mine = 0;
#pragma unroll
      for (u32 q = 1; q <= NL; ++q) {
        const unsigned long long mq = __ballot(b == q);
        if (b == q) mine = mq;
        if (lane == 0) wcnt[q - 1][wv] = u32(__popcll(mq));
      }
      __syncthreads();
      if (t < NL) {
        const u32 tot = wcnt[t][0] + wcnt[t][1] + wcnt[t][2] + wcnt[t][3];
        u32 base = tot ? atomicAdd(&counts[t + 1], tot) : 0;
        for (u32 w = 0; w < 4; ++w) {
          woff[t][w] = base;
          base += wcnt[t][w];
        }
      }
      __syncthreads();
      if (b >= 1 && b <= NL) lists[b * cap + woff[b - 1][wv] + u32(__popcll(mine & below))] = v;
    }
    if (b == 0) {
      // row-order sums per distinct key in registers (static indices only),
      // so the kernel needs no LDS and keeps more wavefronts resident
      const u64 off = g.row_off[v];
      const u32 cur = pol.cur_key(v);
      u32 rk[SMALL_MAX2];
      double rw[SMALL_MAX2];
      u32 nk = 0;
      for (u32 j = 0; j < deg; ++j) {
        const double w = g.w[off + j];
        u32 key;
        int mult;
        if (!pol.template accept_e<false>(v, g.col, off + j, key, mult)) continue;
        bool found = false;
#pragma unroll
        for (u32 q = 0; q < SMALL_MAX2; ++q)
          if (q < nk && rk[q] == key) {
            rw[q] += w;
            if (mult == 2) rw[q] += w;
            found = true;
          }
        if (!found) {
#pragma unroll
          for (u32 q = 0; q < SMALL_MAX2; ++q)
            if (q == nk) {
              rk[q] = key;
              rw[q] = w;
              if (mult == 2) rw[q] += w;
            }
          ++nk;
        }
      }
      vals[2 * kBuckets] += nk;
      double wcur = 0.0;
#pragma unroll
      for (u32 q = 0; q < SMALL_MAX2; ++q)
        if (q < nk && rk[q] == cur) wcur = rw[q];
      const double kv = g.k[v], sc = pol.cur_total(cur), tt = tm_of(g, v);
#pragma unroll
      for (u32 q = 0; q < SMALL_MAX2; ++q) {
        if (q >= nk) continue;
        const u32 key = rk[q];
        if (key == cur || !pol.ok(key)) continue;
        const double gn = move_gain(rw[q], wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
        if (gn > 0.0) better(bg, best, gn, key);
      }
      finish_choice(pol, v, cur, wcur, kv, tt, bg, best);
    }
    {
      // the candidates: one device atomic per workgroup pass
      __shared__ u32 ccnt[4], coff[4];
      const u32 lane = t & 63, wv = t >> 6;
      const bool want = bg > 0.0;
      const unsigned long long m = __ballot(want);
      if (lane == 0) ccnt[wv] = u32(__popcll(m));
      __syncthreads();
      if (t == 0) {
        const u32 tot = ccnt[0] + ccnt[1] + ccnt[2] + ccnt[3];
        u32 base = tot ? atomicAdd(out.cnt, tot) : 0;
        for (u32 w = 0; w < 4; ++w) {
          coff[w] = base;
          base += ccnt[w];
        }
      }
      __syncthreads();
      if (want) {
        const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const u32 p = coff[wv] + u32(__popcll(m & below));
        out.node[p] = v;
        out.to[p] = best;
        out.gain[p] = bg;
      }
    }
  }
  u32 idx[2 * kBuckets + 1];
#pragma unroll
  for (u32 q = 0; q < 2 * kBuckets; ++q) idx[q] = q;
  idx[2 * kBuckets] = 2 * kBuckets;  // the small bucket's candidates: st[2 kBuckets + 0]
  block_add_multi(st, idx, vals);
}

// Proposals are staged in LDS and appended with one device atomic per flush.
constexpr u32 kStage = 256;
__device__ void flush_stage(u32* sn, u32* st, double* sg, u32& scnt, u32& sbase, CandOut out) {
  __syncthreads();
  const u32 c = scnt;
  if (c) {
    if (threadIdx.x == 0) sbase = atomicAdd(out.cnt, c);
    __syncthreads();
    const u32 b = sbase;
    for (u32 i = threadIdx.x; i < c; i += blockDim.x) {
      out.node[b + i] = sn[i];
      out.to[b + i] = st[i];
      out.gain[b + i] = sg[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) scnt = 0;
  }
  __syncthreads();
}

// G lanes per vertex (a wavefront or half of one), an LDS open-addressing
// table of up to CAP slots per vertex (a power of two >= 2 deg, so only the
// slots a vertex can use are cleared); 256 / G vertices per workgroup pass
template <int G>
__device__ __forceinline__ void group_argmax(double& g, u32& k) {
  for (int o = G / 2; o > 0; o >>= 1) {
    double og = __shfl_xor(g, o, 64);
    u32 ok = __shfl_xor(k, o, 64);
    better(g, k, og, ok);
  }
}

template <class P, u32 G, u32 CAP>
__global__ __launch_bounds__(256) void k_prop2_wave(P pol, Csr g, const u32* list, const u32* cnt_dev, CandOut out,
                                                     unsigned long long* cst) {
  constexpr u32 NG = 256 / G;
  __shared__ u32 hk[NG * CAP];
  __shared__ double hv[NG * CAP];
  __shared__ u32 stn[kStage], stt[kStage];
  __shared__ double stg[kStage];
  __shared__ u32 scnt, sbase;
  if (threadIdx.x == 0) scnt = 0;
  const u32 cnt = *cnt_dev;
  const u32 lane = threadIdx.x % G, grp = threadIdx.x / G;
  u32* K = hk + grp * CAP;
  double* V = hv + grp * CAP;
  unsigned long long ncand = 0;
  for (u64 base = u64(blockIdx.x) * NG; base < cnt; base += u64(gridDim.x) * NG) {
    const u64 id = base + grp;
    const bool has = id < cnt;
    u32 v = 0, cur = 0, deg = 0, cap = 64;
    u64 off = 0;
    if (has) {
      v = list[id];
      off = g.row_off[v];
      deg = g.row_len[v];
      cur = pol.cur_key(v);
      while (cap < 2 * deg) cap <<= 1;
      for (u32 q = lane; q < cap; q += G) {
        K[q] = kNone;
        V[q] = 0.0;
      }
    }
    __syncthreads();
    if (has) {
      // four row entries per lane in flight: their gathers are independent
      for (u32 j0 = lane; j0 < deg; j0 += 4 * G) {
        u32 kk[4];
        double ww[4];
        int mm[4];
        bool ok[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const u32 j = j0 + r * G;
          ok[r] = j < deg;
          ww[r] = ok[r] ? ld_row(g.w + off + j) : 0.0;
          ok[r] = ok[r] && pol.accept_e(v, g.col, off + j, kk[r], mm[r]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (!ok[r]) continue;
          const u32 h = tab_insert(K, cap, kk[r]);
          atomicAdd(&V[h], ww[r]);
          if (mm[r] == 2) atomicAdd(&V[h], ww[r]);
        }
      }
    }
    __syncthreads();
    double bg = 0.0, wcur = 0.0;
    u32 best = kNone;
    if (has) {
      wcur = tab_lookup(K, V, cap, cur);
      const double kv = g.k[v], sc = pol.cur_total(cur), tt = tm_of(g, v);
      for (u32 q = lane; q < cap; q += G) {
        u32 key = K[q];
        if (key == kNone) continue;
        ++ncand;
        if (key == cur || !pol.ok(key)) continue;
        double gn = move_gain(V[q], wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
        if (gn > 0.0) better(bg, best, gn, key);
      }
    }
    group_argmax<G>(bg, best);  // every lane of the wavefront takes part in the shuffles
    if (has && lane == 0) {
      finish_choice(pol, v, cur, wcur, g.k[v], tm_of(g, v), bg, best);
      if (bg > 0.0) {
        const u32 p = atomicAdd(&scnt, 1u);  // LDS
        stn[p] = v;
        stt[p] = best;
        stg[p] = bg;
      }
    }
    __syncthreads();
    if (scnt > kStage - NG) flush_stage(stn, stt, stg, scnt, sbase, out);
  }
  flush_stage(stn, stt, stg, scnt, sbase, out);
  block_add_u64(cst, ncand);
}

// one workgroup per vertex of degree <= kLdsDeg, the table in LDS
template <class P>
__global__ __launch_bounds__(1024) void k_prop2_lds(P pol, Csr g, const u32* list, const u32* cnt_dev, CandOut out,
                                                     unsigned long long* cst) {
  __shared__ u32 K[kLdsCap];
  __shared__ double V[kLdsCap];
  const u32 cnt = *cnt_dev;
  unsigned long long ncand = 0;
  for (u32 id = blockIdx.x; id < cnt; id += gridDim.x) {
    const u32 v = list[id];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const u32 cur = pol.cur_key(v);
    const u32 cap = big_cap(deg);
    for (u32 q = threadIdx.x; q < cap; q += blockDim.x) {
      K[q] = kNone;
      V[q] = 0.0;
    }
    __syncthreads();
    for (u32 j = threadIdx.x; j < deg; j += blockDim.x) {
      u32 key;
      int mult;
      if (!pol.accept_e(v, g.col, off + j, key, mult)) continue;
      const double w = ld_row(g.w + off + j);
      const u32 h = tab_insert(K, cap, key);
      atomicAdd(&V[h], w);
      if (mult == 2) atomicAdd(&V[h], w);
    }
    __syncthreads();
    const double wcur = tab_lookup(K, V, cap, cur);
    const double kv = g.k[v], sc = pol.cur_total(cur), tt = tm_of(g, v);
    double bg = 0.0;
    u32 best = kNone;
    for (u32 q = threadIdx.x; q < cap; q += blockDim.x) {
      const u32 key = K[q];
      if (key == kNone) continue;
      ++ncand;
      if (key == cur || !pol.ok(key)) continue;
      const double gn = move_gain(V[q], wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
      if (gn > 0.0) better(bg, best, gn, key);
    }
    block_argmax(bg, best);
    if (threadIdx.x == 0) {
      finish_choice(pol, v, cur, wcur, kv, tt, bg, best);
      if (bg > 0.0) {
        const u32 p = atomicAdd(out.cnt, 1u);
        out.node[p] = v;
        out.to[p] = best;
        out.gain[p] = bg;
      }
    }
    __syncthreads();
  }
  block_add_u64(cst, ncand);
}

// Hubs, first try: one workgroup per vertex with an LDS table of S slots.  A
// hub's neighbours mostly fall in few communities (R-MAT scale 24: a median of
// about 120 distinct ones per hub, at most a few thousand), so the table rarely
// fills; a vertex whose keys do not fit (a key finds no slot within kHubProbe
// probes) is appended to the overflow list for k_prop2_big instead.
constexpr u32 kHubProbe = 64;

__device__ __forceinline__ u32 lds_try_insert(u32* K, u32 mask, u32 key) {
  u32 h = hash32(key) & mask;
  for (u32 t = 0; t < kHubProbe; ++t) {
    const u32 old = atomicCAS(&K[h], kNone, key);
    if (old == kNone || old == key) return h;
    h = (h + 1) & mask;
  }
  return kNone;
}
__device__ __forceinline__ double lds_find(const u32* K, const double* V, u32 mask, u32 key) {
  u32 h = hash32(key) & mask;
  for (u32 t = 0; t < kHubProbe; ++t) {
    const u32 x = K[h];
    if (x == key) return V[h];
    if (x == kNone) return 0.0;
    h = (h + 1) & mask;
  }
  return 0.0;
}

// Hub table of vertex v written from its workgroup's LDS table (K, V of S
// slots).  cap = the power of two >= 2 c + 16 for c distinct keys, so that new
// keys from the neighbours' moves fit; when cap <= S / 2 the keys are rehashed
// into the LDS scratch (K2, V2) first, else the LDS table is the table
// (cap = S, fewer slots than the row has entries).  The HBM table takes plain
// stores only (the kernels that read it come later); pool space comes in
// chunks of up to four tables per workgroup (one device atomic per chunk); no
// table when the pool is spent.
constexpr u64 kHtChunk = 8192;
template <u32 S>
__device__ __forceinline__ void build_hub_table(const HubTables& ht, u32 v, const u32* K, const double* V, u32* K2,
                                                double* V2, u32 c, u64& chunk_at, u64& chunk_end) {
  __shared__ u64 s_off;
  u32 cap = 32;
  while (cap < 2 * c + 16) cap <<= 1;
  if (cap > S / 2) cap = S;
  if (threadIdx.x == 0) {
    if (chunk_at + cap > chunk_end) {
      const u64 want = cap >= kHtChunk / 4 ? cap : 4 * cap;
      chunk_at = atomicAdd(ht.used, (unsigned long long)want);
      chunk_end = chunk_at + want;
      if (chunk_end > ht.pool) chunk_at = chunk_end = ht.pool;  // spent
    }
    s_off = chunk_at + cap <= chunk_end ? chunk_at : ~0ull;
    if (s_off != ~0ull) chunk_at += cap;
  }
  if (cap < S) {
    for (u32 q = threadIdx.x; q < cap; q += blockDim.x) {
      K2[q] = kNone;
      V2[q] = 0.0;
    }
    __syncthreads();
    for (u32 q = threadIdx.x; q < S; q += blockDim.x) {
      const u32 key = K[q];
      if (key == kNone) continue;
      V2[tab_insert(K2, cap, key)] = V[q];
    }
    K = K2;
    V = V2;
  }
  __syncthreads();
  const u64 off = s_off;
  if (off == ~0ull) return;
  for (u32 q = threadIdx.x; q < cap; q += blockDim.x) {
    ht.slot[off + q] = HtSlot{K[q], 0u, V[q]};
  }
  if (threadIdx.x == 0) {
    ht.loc[v] = off | (u64(__ffs(int(cap)) - 1) << 56);
    atomicOr(&ht.bits[v >> 5], 1u << (v & 31));
  }
}

template <class P, u32 T, u32 S>
__global__ __launch_bounds__(T) void k_prop2_hub(P pol, Csr g, const u32* list, const u32* cnt_dev, u32* ovf,
                                                 u32* ovf_cnt, CandOut out, unsigned long long* cst,
                                                 unsigned long long* tst) {
  // hub tables for the rows above the wavefront bucket
  constexpr bool kTab = P::kTables;
  constexpr u32 S2 = kTab ? S / 2 : 1;
  __shared__ u32 K[S];
  __shared__ double V[S];
  __shared__ u32 K2[S2];  // the hub table rehashed to its own size
  __shared__ double V2[S2];
  __shared__ u32 full;
  __shared__ u32 csum[16];
  u64 chunk_at = 0, chunk_end = 0;  // this workgroup's pool chunk (thread 0)
  const u32 cnt = *cnt_dev;
  unsigned long long ncand = 0, nv = 0, ne = 0;
  for (u32 id = blockIdx.x; id < cnt; id += gridDim.x) {
    const u32 v = list[id];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const u32 cur = pol.cur_key(v);
    for (u32 q = threadIdx.x; q < S; q += T) {
      K[q] = kNone;
      V[q] = 0.0;
    }
    if (threadIdx.x == 0) full = 0;
    __syncthreads();
    // four row entries per thread in flight: the gathers are independent
    for (u32 j0 = threadIdx.x; j0 < deg; j0 += 4 * T) {
      if (full) break;
      u32 kk[4];
      double ww[4];
      int mm[4];
      bool ok[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const u32 j = j0 + r * T;
        ok[r] = j < deg;
        ww[r] = ok[r] ? ld_row(g.w + off + j) : 0.0;
        ok[r] = ok[r] && pol.accept_e(v, g.col, off + j, kk[r], mm[r]);
      }
      bool spilled = false;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!ok[r] || spilled) continue;
        const u32 h = lds_try_insert(K, S - 1, kk[r]);
        spilled = h == kNone;
        if (!spilled) {
          atomicAdd(&V[h], ww[r]);
          if (mm[r] == 2) atomicAdd(&V[h], ww[r]);
        }
      }
      if (spilled) full = 1;
    }
    __syncthreads();
    if (full) {  // uniform: nothing writes it between the barrier above and the one below
      if (threadIdx.x == 0) ovf[atomicAdd(ovf_cnt, 1u)] = v;
      __syncthreads();
      continue;
    }
    const double wcur = lds_find(K, V, S - 1, cur);
    const double kv = g.k[v], sc = pol.cur_total(cur), tt = tm_of(g, v);
    double bg = 0.0;
    u32 best = kNone, lc = 0;
    for (u32 q = threadIdx.x; q < S; q += T) {
      const u32 key = K[q];
      if (key == kNone) continue;
      ++lc;
      if (key == cur || !pol.ok(key)) continue;
      const double gn = move_gain(V[q], wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
      if (gn > 0.0) better(bg, best, gn, key);
    }
    ncand += lc;
    block_argmax(bg, best);
    if (threadIdx.x == 0) {
      nv += 1;
      ne += deg;
      finish_choice(pol, v, cur, wcur, kv, tt, bg, best);
      if (bg > 0.0) {
        const u32 p = atomicAdd(out.cnt, 1u);
        out.node[p] = v;
        out.to[p] = best;
        out.gain[p] = bg;
      }
    }
    if constexpr (kTab) {
      if (pol.ht.bits) {
        // distinct keys of the row, then the table
        u32 x = lc;
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) csum[threadIdx.x >> 6] = x;
        __syncthreads();
        u32 c = 0;
        for (u32 i = 0; i < T / 64; ++i) c += csum[i];
        build_hub_table<S>(pol.ht, v, K, V, K2, V2, c, chunk_at, chunk_end);
      }
    }
    __syncthreads();
  }
  block_add_u64(cst, ncand);
  const u32 idx[3] = {0, 1, 2};
  unsigned long long vals[3] = {nv, ne, ncand};
  block_add_multi(tst, idx, vals);
}

// Vertices with a valid hub table: one wavefront per vertex scans the table's
// slots (the live keys are its neighbour communities) instead of the row.
template <class P>
__global__ __launch_bounds__(256) void k_prop2_ht(P pol, Csr g, const u32* list, const u32* cnt_dev, CandOut out,
                                                   unsigned long long* hst) {
  __shared__ u32 stn[kStage], stt[kStage];
  __shared__ double stg[kStage];
  __shared__ u32 scnt, sbase;
  if (threadIdx.x == 0) scnt = 0;
  const u32 cnt = *cnt_dev;
  const u32 lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  unsigned long long nv = 0, nslots = 0, ncand = 0;
  __syncthreads();
  for (u64 base = u64(blockIdx.x) * 4; base < cnt; base += u64(gridDim.x) * 4) {
    const u64 id = base + grp;
    const bool has = id < cnt;  // uniform over the wavefront
    double bg = 0.0, wcur = 0.0, kv = 0.0, sc = 0.0, tt = 0.0;
    u32 best = kNone, v = 0, cur = 0;
    if (has) {
      v = list[id];
      cur = pol.cur_key(v);
      u64 off;
      u32 cap;
      pol.ht.where(v, off, cap);
      const HtSlot* T = pol.ht.slot + off;
      for (u32 q = lane; q < cap; q += 64)
        if (T[q].key == cur) wcur = T[q].w;  // 0 when v has no neighbour left there
      for (int o = 32; o > 0; o >>= 1) wcur += __shfl_xor(wcur, o, 64);  // one lane at most holds it
      kv = g.k[v];
      sc = pol.cur_total(cur);
      tt = tm_of(g, v);
      for (u32 q = lane; q < cap; q += 64) {
        const HtSlot sl = T[q];
        const u32 key = sl.key;
        const double wq = sl.w;
        if (key == kNone || wq == 0.0) continue;
        ++ncand;
        if (key == cur) continue;
        const double gn = move_gain(wq, wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
        if (gn > 0.0) better(bg, best, gn, key);
      }
      if (lane == 0) {
        nv += 1;
        nslots += cap;
      }
    }
    wave_argmax(bg, best);
    if (has && lane == 0) {
      finish_choice(pol, v, cur, wcur, kv, tt, bg, best);
      if (bg > 0.0) {
        const u32 p = atomicAdd(&scnt, 1u);  // LDS
        stn[p] = v;
        stt[p] = best;
        stg[p] = bg;
      }
    }
    __syncthreads();
    if (scnt > kStage - 4) flush_stage(stn, stt, stg, scnt, sbase, out);
  }
  flush_stage(stn, stt, stg, scnt, sbase, out);
  const u32 idx[3] = {0, 1, 2};
  unsigned long long vals[3] = {nv, nslots, ncand};
  block_add_multi(hst, idx, vals);
}

// A neighbour v of hub u moved from community `from` to `to` (row weight w):
// keep u's table current.  The lookup of `from` (a key since before this
// kernel) takes plain loads; a key that finds no slot drops the table (u is
// then evaluated from its row, which rebuilds it).
__device__ void ht_move(const HubTables& ht, u32 u, u32 from, u32 to, double w) {
  if (!ht.valid(u)) return;
  u64 off;
  u32 cap;
  ht.where(u, off, cap);
  HtSlot* T = ht.slot + off;
  u32 h = hash32(from) & (cap - 1);
  bool found = false;
  for (u32 t = 0; t < cap; ++t) {
    const u32 k = T[h].key;
    if (k == from) {
      found = true;
      break;
    }
    if (k == kNone) break;
    h = (h + 1) & (cap - 1);
  }
  if (!found) {  // cannot happen for a table built from the current row
    atomicAnd(&ht.bits[u >> 5], ~(1u << (u & 31)));
    return;
  }
  atomicAdd(&T[h].w, -w);
  // `to`: an existing key is found by plain loads (no returning atomic on the
  // common path); an empty-looking slot is claimed with a compare-and-swap,
  // which also settles a load that missed a key inserted during this kernel
  h = hash32(to) & (cap - 1);
  for (u32 t = 0; t < cap; ++t) {
    u32 k = T[h].key;
    if (k == kNone) k = atomicCAS(&T[h].key, kNone, to);
    if (k == kNone || k == to) {
      atomicAdd(&T[h].w, w);
      return;
    }
    h = (h + 1) & (cap - 1);
  }
  atomicAnd(&ht.bits[u >> 5], ~(1u << (u & 31)));
}

// G lanes per vertex with a small LDS table of at most CAP slots: the
// distinct neighbour communities of a vertex are usually far fewer than its
// degree (R-MAT scale 24 after a batch: at most 36 for degrees 17-128, 56 for
// 129-512), and small tables give more resident wavefronts and less to clear
// and scan.  A vertex whose keys do not fit is appended to the overflow list
// for the degree-sized tables instead.
template <u32 G, u32 CAP, u32 BS>
struct TabShared {
  u32 hk[BS / G * CAP];
  double hv[BS / G * CAP];
  u32 gfull[BS / G];
  u32 stn[kStage], stt[kStage];
  double stg[kStage];
  u32 scnt, sbase;
};
// a thread's share of the work counters of tab_body calls, added to the
// device counters once (the persistent rounds add them when they end)
struct TabTally {
  unsigned long long nv = 0, ne = 0, ncand = 0;
};
// G lanes per vertex in workgroups of BS threads; tst2 (or null) takes the
// evaluated vertices and row entries as well; with tally the counters are
// added there instead
template <class P, u32 G, u32 CAP, u32 BS>
__device__ __forceinline__ void tab_body(const P& pol, const Csr& g, const u32* list, const u32* cnt_dev, u32* ovf,
                                         u32* ovf_cnt, const CandOut& out, unsigned long long* cst,
                                         unsigned long long* tst, TabShared<G, CAP, BS>& sh,
                                         unsigned long long* tst2 = nullptr, TabTally* tally = nullptr) {
  constexpr u32 NG = BS / G;
  u32* hk = sh.hk;
  double* hv = sh.hv;
  u32* gfull = sh.gfull;
  u32* stn = sh.stn;
  u32* stt = sh.stt;
  double* stg = sh.stg;
  u32& scnt = sh.scnt;
  u32& sbase = sh.sbase;
  if (threadIdx.x == 0) scnt = 0;
  const u32 cnt = *cnt_dev;
  const u32 lane = threadIdx.x % G, grp = threadIdx.x / G;
  u32* K = hk + grp * CAP;
  double* V = hv + grp * CAP;
  unsigned long long ncand = 0, nv = 0, ne = 0;
  // the next pass's vertex and row header are loaded one pass ahead, so
  // their latency overlaps this pass's gathers
  const u64 step = u64(gridDim.x) * NG;
  u32 v_nx = 0, deg_nx = 0;
  u64 off_nx = 0;
  {
    const u64 id0 = u64(blockIdx.x) * NG + grp;
    if (id0 < cnt) {
      v_nx = list[id0];
      off_nx = g.row_off[v_nx];
      deg_nx = g.row_len[v_nx];
    }
  }
  for (u64 base = u64(blockIdx.x) * NG; base < cnt; base += step) {
    const u64 id = base + grp;
    const bool has = id < cnt;
    u32 v = 0, cur = 0, deg = 0, cap = 16;
    u64 off = 0;
    if (has) {
      v = v_nx;
      off = off_nx;
      deg = deg_nx;
      if (id + step < cnt) {
        v_nx = list[id + step];
        off_nx = g.row_off[v_nx];
        deg_nx = g.row_len[v_nx];
      }
      cur = pol.cur_key(v);
      while (cap < 2 * deg && cap < CAP) cap <<= 1;
      for (u32 q = lane; q < cap; q += G) {
        K[q] = kNone;
        V[q] = 0.0;
      }
    }
    if (lane == 0) gfull[grp] = 0;
    __syncthreads();
    if (has) {
      // four row entries per lane in flight: their gathers are independent
      bool spilled = false;
      for (u32 j0 = lane; j0 < deg && !spilled; j0 += 4 * G) {
        u32 kk[4];
        double ww[4];
        int mm[4];
        bool ok[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const u32 j = j0 + r * G;
          ok[r] = j < deg;
          ww[r] = ok[r] ? ld_row(g.w + off + j) : 0.0;
          ok[r] = ok[r] && pol.accept_e(v, g.col, off + j, kk[r], mm[r]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (!ok[r] || spilled) continue;
          const u32 h = lds_try_insert(K, cap - 1, kk[r]);
          spilled = h == kNone;
          if (!spilled) {
            atomicAdd(&V[h], ww[r]);
            if (mm[r] == 2) atomicAdd(&V[h], ww[r]);
          }
        }
      }
      if (spilled) gfull[grp] = 1;
    }
    __syncthreads();
    const bool spill = has && gfull[grp] != 0;  // written before the barrier above, cleared after the next
    if (spill && lane == 0) ovf[atomicAdd(ovf_cnt, 1u)] = v;
    double bg = 0.0, wcur = 0.0;
    u32 best = kNone;
    if (has && !spill) {
      wcur = lds_find(K, V, cap - 1, cur);
      const double kv = g.k[v], sc = pol.cur_total(cur), tt = tm_of(g, v);
      for (u32 q = lane; q < cap; q += G) {
        const u32 key = K[q];
        if (key == kNone) continue;
        ++ncand;
        if (key == cur || !pol.ok(key)) continue;
        const double gn = move_gain(V[q], wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
        if (gn > 0.0) better(bg, best, gn, key);
      }
    }
    group_argmax<G>(bg, best);  // every lane of the wavefront takes part in the shuffles
    if (has && !spill && lane == 0) {
      nv += 1;
      ne += deg;
      finish_choice(pol, v, cur, wcur, g.k[v], tm_of(g, v), bg, best);
      if (bg > 0.0) {
        const u32 p = atomicAdd(&scnt, 1u);  // LDS
        stn[p] = v;
        stt[p] = best;
        stg[p] = bg;
      }
    }
    __syncthreads();
    if (scnt > kStage - NG) flush_stage(stn, stt, stg, scnt, sbase, out);
  }
  flush_stage(stn, stt, stg, scnt, sbase, out);
  if (tally) {
    tally->nv += nv;
    tally->ne += ne;
    tally->ncand += ncand;
    return;
  }
  block_add_u64(cst, ncand);
  const u32 idx[3] = {0, 1, 2};
  unsigned long long vals[3] = {nv, ne, ncand};
  if (tst2) {
    const u32 idx2[2] = {0, kBuckets};
    unsigned long long vals2[2] = {nv, ne};
    block_add_multi(tst2, idx2, vals2);
  }
  block_add_multi(tst, idx, vals);
}
template <class P, u32 G, u32 CAP>
__global__ __launch_bounds__(256) void k_prop2_tab(P pol, Csr g, const u32* list, const u32* cnt_dev, u32* ovf,
                                                    u32* ovf_cnt, CandOut out, unsigned long long* cst,
                                                    unsigned long long* tst) {
  __shared__ TabShared<G, CAP, 256> sh;
  tab_body<P, G, CAP, 256>(pol, g, list, cnt_dev, ovf, ovf_cnt, out, cst, tst, sh);
}

// Hubs whose keys overflowed the LDS table: one workgroup per vertex; each
// resident workgroup owns an HBM table of tcap slots (cleared per vertex: coalesced, cheaper than tracking the
// used slots, whose clears and scans are random accesses)
// An LDS table in front of the HBM one absorbs the heavy keys: a hub's
// neighbours often share a few big communities, and device atomics on one
// slot serialise.  Keys that find no LDS slot within kCacheProbe probes go to
// the HBM table; the LDS entries are flushed into it after the row.
constexpr u32 kCacheSlots = 4096;
constexpr u32 kCacheProbe = 8;

template <class P>
__global__ __launch_bounds__(1024) void k_prop2_big(P pol, Csr g, const u32* list, const u32* cnt_dev, u32* TK,
                                                     double* TV, u32 tcap, CandOut out, unsigned long long* cst) {
  __shared__ u32 LK[kCacheSlots];
  __shared__ double LV[kCacheSlots];
  const u32 cnt = *cnt_dev;
  unsigned long long ncand = 0;
  u32* K = TK + u64(blockIdx.x) * tcap;
  double* V = TV + u64(blockIdx.x) * tcap;
  for (u32 id = blockIdx.x; id < cnt; id += gridDim.x) {
    const u32 v = list[id];
    const u64 off = g.row_off[v];
    const u32 deg = g.row_len[v];
    const u32 cur = pol.cur_key(v);
    const u32 cap = big_cap(deg);
    for (u32 q = threadIdx.x; q < cap; q += blockDim.x) {
      K[q] = kNone;
      V[q] = 0.0;
    }
    for (u32 q = threadIdx.x; q < kCacheSlots; q += blockDim.x) {
      LK[q] = kNone;
      LV[q] = 0.0;
    }
    __threadfence();
    __syncthreads();
    // four row entries per thread in flight: the gathers are independent
    for (u32 j0 = threadIdx.x; j0 < deg; j0 += 4 * blockDim.x) {
      u32 kk[4];
      double ww[4];
      int mm[4];
      bool ok[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const u32 j = j0 + r * blockDim.x;
        ok[r] = j < deg;
        ww[r] = ok[r] ? ld_row(g.w + off + j) : 0.0;
        ok[r] = ok[r] && pol.accept_e(v, g.col, off + j, kk[r], mm[r]);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!ok[r]) continue;
        u32 h = hash32(kk[r]) & (kCacheSlots - 1);
        bool cached = false;
        for (u32 t = 0; t < kCacheProbe; ++t) {
          const u32 old = atomicCAS(&LK[h], kNone, kk[r]);
          if (old == kNone || old == kk[r]) {
            cached = true;
            break;
          }
          h = (h + 1) & (kCacheSlots - 1);
        }
        if (cached) {
          atomicAdd(&LV[h], ww[r]);
          if (mm[r] == 2) atomicAdd(&LV[h], ww[r]);
        } else {
          const u32 hh = hbm_insert(K, cap, kk[r]);
          atomicAdd(&V[hh], ww[r]);
          if (mm[r] == 2) atomicAdd(&V[hh], ww[r]);
        }
      }
    }
    __syncthreads();
    for (u32 q = threadIdx.x; q < kCacheSlots; q += blockDim.x) {
      const u32 key = LK[q];
      if (key == kNone) continue;
      const u32 hh = hbm_insert(K, cap, key);
      atomicAdd(&V[hh], LV[q]);
    }
    __threadfence();
    __syncthreads();
    const double wcur = tab_lookup(K, V, cap, cur);
    const double kv = g.k[v], sc = pol.cur_total(cur), tt = tm_of(g, v);
    double bg = 0.0;
    u32 best = kNone;
    for (u32 q = threadIdx.x; q < cap; q += blockDim.x) {
      const u32 key = K[q];
      if (key == kNone) continue;
      ++ncand;
      if (key == cur || !pol.ok(key)) continue;
      const double gn = move_gain(V[q], wcur, tt, pol.gamma, kv, sc, pol.cand_total(key));
      if (gn > 0.0) better(bg, best, gn, key);
    }
    block_argmax(bg, best);
    if (threadIdx.x == 0) {
      finish_choice(pol, v, cur, wcur, kv, tt, bg, best);
      if (bg > 0.0) {
        const u32 p = atomicAdd(out.cnt, 1u);
        out.node[p] = v;
        out.to[p] = best;
        out.gain[p] = bg;
      }
    }
    __syncthreads();
  }
  block_add_u64(cst, ncand);
}

// ------------------------------------------------------------------ decoupling (one workgroup)
// (gain desc, node asc) as a sortable pair: key = ~bits(gain) (gains are > 0)
__device__ __forceinline__ bool pri_less(u64 ka, u32 na, u64 kb, u32 nb) { return ka < kb || (ka == kb && na < nb); }

// the decoupling workgroup's LDS (k_dec_small, and the persistent rounds)
struct DecShared {
  u64 skey[kDecSmall];
  u32 snode[kDecSmall];
  u32 sidx[kDecSmall];
  u32 sfrom[kDecSmall];
  u32 sto[kDecSmall];
  u8 sst[kDecSmall];
  u8 skd[kDecSmall];  // the accepted move's vertex moved in an earlier round of this level
  u32 mAto[kDecHash], mAfrom[kDecHash];
  u16 fslot[kDecLds], tslot[kDecLds];
  u32 und, nacc;
  u64 wsum[16], mtotal;
};
// one 1,024-thread workgroup
__device__ __forceinline__ void dec_small_body(const MoveCtx& m, const CandOut& cand, u32* minUto, u32* minUfrom,
                                               u32* minAto, u32* minAfrom, u32* mark, u32* next, u32* next_cnt,
                                               u32* status, unsigned long long* dst, u32* cur_cnt,
                                               u32* bucket_counts, const FlatMarks& flat, u32* round_cand,
                                               DecShared& sh) {
  u64* skey = sh.skey;
  u32* snode = sh.snode;
  u32* sidx = sh.sidx;
  u32* sfrom = sh.sfrom;
  u32* sto = sh.sto;
  u8* sst = sh.sst;
  u8* skd = sh.skd;
  u32* mAto = sh.mAto;
  u32* mAfrom = sh.mAfrom;
  u16* fslot = sh.fslot;
  u16* tslot = sh.tslot;
  u32& und = sh.und;
  u32& nacc = sh.nacc;
  u64* wsum = sh.wsum;
  u64& mtotal = sh.mtotal;
  const u32 c = *cand.cnt;
  const u32 tid = threadIdx.x;
  if (tid == 0) *round_cand = c;  // read by the host with the round's other counters
  if (c > kDecSmall) {
    if (tid == 0) {
      status[0] = 1;  // multi-kernel path: it keeps the candidates and their count
      *flat.c = 0;
      if (m.ck) *flat.ck_c = 0;
      // the round's list and bucket counts are spent: a round enqueued after
      // this one finds them empty
      *cur_cnt = 0;
      clear_bucket_counts(bucket_counts);
    }
    return;
  }
  if (c <= 64) {
    // one wavefront: register bitonic sort on (gain desc, node asc), then the
    // greedy acceptance of throughput.rs:217-233 in priority order by ballots
    if (tid < 64) {
      const u32 lane = tid;
      u64 key = ~0ull;
      u32 node = kNone, to = 0, from = 0;
      if (lane < c) {
        key = ~(u64)__double_as_longlong(cand.gain[lane]);
        node = cand.node[lane];
        to = cand.to[lane];
        from = m.f[node];
      }
      for (u32 kk = 2; kk <= 64; kk <<= 1) {
        for (u32 j = kk >> 1; j > 0; j >>= 1) {
          const u32 klo = __shfl_xor(u32(key), j, 64), khi = __shfl_xor(u32(key >> 32), j, 64);
          const u64 ok = (u64(khi) << 32) | klo;
          const u32 on = __shfl_xor(node, j, 64), ot = __shfl_xor(to, j, 64), of = __shfl_xor(from, j, 64);
          const bool keep_min = ((lane & j) == 0) == ((lane & kk) == 0);
          const bool take = keep_min ? pri_less(ok, on, key, node) : pri_less(key, node, ok, on);
          if (take) {
            key = ok;
            node = on;
            to = ot;
            from = of;
          }
        }
      }
      bool acc = false;
      for (u32 p = 0; p < c; ++p) {
        const u32 fp = __shfl(from, p, 64), tp = __shfl(to, p, 64);
        const bool conf = lane < p && acc && (to == fp || from == tp);
        const unsigned long long bm = __ballot(conf);
        if (lane == p) acc = bm == 0;
      }
      if (lane < c) {
        snode[lane] = node;
        sto[lane] = to;
        sfrom[lane] = from;
        sst[lane] = acc ? kAcc : kRej;
        skd[lane] = acc ? m.B[node] : u8(0);  // moved in an earlier round of this level
      }
      if (acc) {
        const double kv = m.k[node];
        st_agent(&m.f[node], to);
        if (m.fs) st_agent(fs_f(m.fs, node), to);
        atomicAdd(&m.tot[from], -kv);
        atomicAdd(&m.tot[to], kv);
        m.B[node] = 1;
      }
      const unsigned long long am = __ballot(acc);
      if (lane == 0) nacc = u32(__popcll(am));
    }
    __syncthreads();
  } else {
    if (c <= 1024) {
      // rank sort on (key, node), one candidate per thread: every thread counts
      // the candidates ordered before its own (nodes are distinct, so ranks are
      // too); all threads read the same entry at a time, an LDS broadcast
      u64 mk = ~0ull;
      u32 mn = kNone;
      if (tid < c) {
        mk = ~(u64)__double_as_longlong(cand.gain[tid]);
        mn = cand.node[tid];
        skey[tid] = mk;
        snode[tid] = mn;
      }
      __syncthreads();
      u32 rank = 0;
      if (tid < c)
        for (u32 j = 0; j < c; ++j) rank += pri_less(skey[j], snode[j], mk, mn) ? 1u : 0u;
      __syncthreads();
      if (tid < c) {
        skey[rank] = mk;
        snode[rank] = mn;
        sidx[rank] = tid;
      }
      __syncthreads();
    } else {
      u32 N = 1;
      while (N < c) N <<= 1;
      for (u32 i = tid; i < N; i += blockDim.x) {
        if (i < c) {
          skey[i] = ~(u64)__double_as_longlong(cand.gain[i]);
          snode[i] = cand.node[i];
          sidx[i] = i;
        } else {
          skey[i] = ~0ull;
          snode[i] = kNone;
          sidx[i] = kNone;
        }
      }
      __syncthreads();
      // bitonic sort of (key, node) ascending
      for (u32 kk = 2; kk <= N; kk <<= 1) {
        for (u32 j = kk >> 1; j > 0; j >>= 1) {
          for (u32 i = tid; i < N; i += blockDim.x) {
            u32 ixj = i ^ j;
            if (ixj > i) {
              bool up = (i & kk) == 0;
              bool gt = pri_less(skey[ixj], snode[ixj], skey[i], snode[i]);
              if (gt == up) {
                u64 tk = skey[i];
                skey[i] = skey[ixj];
                skey[ixj] = tk;
                u32 tn = snode[i];
                snode[i] = snode[ixj];
                snode[ixj] = tn;
                u32 ti = sidx[i];
                sidx[i] = sidx[ixj];
                sidx[ixj] = ti;
              }
            }
          }
          __syncthreads();
        }
      }
    }
    for (u32 p = tid; p < c; p += blockDim.x) {
      u32 v = snode[p];
      sfrom[p] = m.f[v];
      sto[p] = cand.to[sidx[p]];
      sst[p] = kUndec;
    }
    __syncthreads();
    // priority rounds: reject if an earlier accepted move entered our source or
    // left our target; wait while an earlier undecided one could
    if (c <= kDecLds) {
      // communities -> slots of an LDS hash table (keys in the dead sort keys,
      // per-slot minima in LDS), so a round touches no device memory
      u32* hk = reinterpret_cast<u32*>(skey);  // kDecHash keys
      u32* mUto = hk + kDecHash;
      u32* mUfrom = sidx;
      for (u32 q = tid; q < kDecHash; q += blockDim.x) hk[q] = kNone;
      __syncthreads();
      for (u32 p = tid; p < c; p += blockDim.x) {
        fslot[p] = u16(lds_insert(hk, kDecHash, sfrom[p]));
        tslot[p] = u16(lds_insert(hk, kDecHash, sto[p]));
      }
      __syncthreads();
      for (int iter = 0; iter < (1 << 20); ++iter) {
        for (u32 q = tid; q < kDecHash; q += blockDim.x) {
          mUto[q] = kNone;
          mUfrom[q] = kNone;
          mAto[q] = kNone;
          mAfrom[q] = kNone;
        }
        if (tid == 0) und = 0;
        __syncthreads();
        for (u32 p = tid; p < c; p += blockDim.x) {
          const u8 st = sst[p];
          if (st == kUndec) {
            atomicMin(&mUto[tslot[p]], p);
            atomicMin(&mUfrom[fslot[p]], p);
          } else if (st == kAcc) {
            atomicMin(&mAto[tslot[p]], p);
            atomicMin(&mAfrom[fslot[p]], p);
          }
        }
        __syncthreads();
        u32 local = 0;
        for (u32 p = tid; p < c; p += blockDim.x) {
          if (sst[p] != kUndec) continue;
          const u32 f = fslot[p], t = tslot[p];
          if (mAto[f] < p || mAfrom[t] < p)
            sst[p] = kRej;
          else if (mUto[f] < p || mUfrom[t] < p)
            ++local;
          else
            sst[p] = kAcc;
        }
        if (local) atomicAdd(&und, local);
        __syncthreads();
        const u32 u = und;
        __syncthreads();
        if (u == 0) break;
      }
    } else {
      // priority rounds: reject if an earlier accepted move entered our source or
      // left our target; wait while an earlier undecided one could
      for (int iter = 0; iter < (1 << 20); ++iter) {
        for (u32 p = tid; p < c; p += blockDim.x) {
          u32 f = sfrom[p], t = sto[p];
          atomicExch(&minUto[f], kNone); atomicExch(&minUto[t], kNone);
          atomicExch(&minUfrom[f], kNone); atomicExch(&minUfrom[t], kNone);
          atomicExch(&minAto[f], kNone); atomicExch(&minAto[t], kNone);
          atomicExch(&minAfrom[f], kNone); atomicExch(&minAfrom[t], kNone);
        }
        if (tid == 0) und = 0;
        __threadfence();
        __syncthreads();
        for (u32 p = tid; p < c; p += blockDim.x) {
          u8 st = sst[p];
          if (st == kUndec) {
            atomicMin(&minUto[sto[p]], p);
            atomicMin(&minUfrom[sfrom[p]], p);
          } else if (st == kAcc) {
            atomicMin(&minAto[sto[p]], p);
            atomicMin(&minAfrom[sfrom[p]], p);
          }
        }
        __threadfence();
        __syncthreads();
        u32 local = 0;
        for (u32 p = tid; p < c; p += blockDim.x) {
          if (sst[p] != kUndec) continue;
          u32 f = sfrom[p], t = sto[p];
          if (atomicAdd(&minAto[f], 0u) < p || atomicAdd(&minAfrom[t], 0u) < p)
            sst[p] = kRej;
          else if (atomicAdd(&minUto[f], 0u) < p || atomicAdd(&minUfrom[t], 0u) < p)
            ++local;
          else
            sst[p] = kAcc;
        }
        if (local) atomicAdd(&und, local);
        __syncthreads();
        const u32 u = und;
        __syncthreads();
        if (u == 0) break;
      }
    }
    // apply (throughput.rs:182-187)
    if (tid == 0) nacc = 0;
    __syncthreads();
    u32 la = 0;
    for (u32 p = tid; p < c; p += blockDim.x) {
      if (sst[p] != kAcc) continue;
      ++la;
      u32 v = snode[p];
      double kv = m.k[v];
      st_agent(&m.f[v], sto[p]);
      if (m.fs) st_agent(fs_f(m.fs, v), sto[p]);
      atomicAdd(&m.tot[sfrom[p]], -kv);
      atomicAdd(&m.tot[sto[p]], kv);
      skd[p] = m.B[v];  // one move per vertex and round: no race on B[v]
      m.B[v] = 1;
    }
    if (la) atomicAdd(&nacc, la);
    __syncthreads();
  }
  // next-active and K marks (throughput.rs:189-199) over the flattened
  // neighbour lists of the accepted moves, so that a hub spreads over the
  // whole block; long lists go to k_marks_flat, which spreads them over the grid
  const u32 lane = tid & 63, wv = tid >> 6;
  u64* pre = skey;  // exclusive degree prefix over positions (the sort keys are dead)
  {
    const u32 base = tid * 4;  // c <= 4 * blockDim.x
    u64 loc[4], sum = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const u32 p = base + r;
      loc[r] = sum;
      if (p < c && sst[p] == kAcc) sum += m.g.row_len[snode[p]];
    }
    u64 x = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const u32 lo = __shfl_up(u32(x), o, 64), hi = __shfl_up(u32(x >> 32), o, 64);
      if (lane >= u32(o)) x += (u64(hi) << 32) | lo;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    u64 before = 0;
    for (u32 i = 0; i < wv; ++i) before += wsum[i];
    const u64 excl = before + x - sum;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (base + r < c) pre[base + r] = excl + loc[r];
    if (tid == blockDim.x - 1) mtotal = before + x;
    __syncthreads();
  }
  const u64 total = mtotal;
  if (total <= kMarksInBlock) {
    for (u64 t0 = 0; t0 < total; t0 += blockDim.x) {
      const u64 t = t0 + tid;
      bool want = false;
      u32 u = 0;
      if (t < total) {
        u32 lo = 0, hi = c;  // pre[lo] <= t < pre[hi], with pre[c] = total
        while (hi - lo > 1) {
          const u32 mid = (lo + hi) >> 1;
          if (pre[mid] <= t) lo = mid;
          else hi = mid;
        }
        const u32 v = snode[lo], cto = sto[lo];
        const u64 e = m.g.row_off[v] + (t - pre[lo]);
        u = ld_row(m.g.col + e);
        if (m.ht.bits && u != v) ht_move(m.ht, u, sfrom[lo], cto, ld_row(m.g.w + e));
        if (skd[lo]) {
          // a repeat move: K marks done; a neighbour already claimed needs no gather
          if (!marked(mark, u) && ld_agent(m.fs ? fs_f(m.fs, u) : &m.f[u]) != cto) want = claim(mark, u);
        } else {
          u32 fu, su;
          if (m.fs) {
            const u64 x = ld_agent64(m.fs + u);
            fu = u32(x);
            su = u32(x >> 32);
          } else {
            fu = ld_agent(&m.f[u]);
            su = m.s[u];
          }
          if (fu != cto) want = claim(mark, u);
          if (su == m.s[v]) {
            m.K[u] = 1;
            m.K[v] = 1;
          }
        }
      }
      const u32 q = wave_append(want, next_cnt);
      if (want) next[q] = u;
    }
    if (tid == 0) *flat.c = 0;
  }
  if (total > kMarksInBlock || m.ck) {
    // the moves for k_marks_flat (flat.c set) and for k_ck_flat (ck_c set)
    for (u32 p = tid; p < c; p += blockDim.x) {
      flat.node[p] = snode[p];
      flat.to[p] = sto[p];
      flat.pre[p] = pre[p];
      if (flat.from) flat.from[p] = sfrom[p];
      if (flat.kd) flat.kd[p] = sst[p] == kAcc ? skd[p] : u8(0);
    }
    if (tid == 0) {
      flat.pre[c] = total;
      if (total > kMarksInBlock) *flat.c = c;
      if (m.ck) *flat.ck_c = c;
    }
  }
  if (tid == 0 && total) atomicAdd(&dst[total <= kMarksInBlock ? DS_DEC_DEG : DS_MULTI_DEG], (unsigned long long)total);
  __syncthreads();
  if (tid == 0) {
    status[0] = 0;
    status[1] = nacc;
    // the next round starts from zero counts without a host-side reset
    *cand.cnt = 0;
    *cur_cnt = 0;
    clear_bucket_counts(bucket_counts);
    atomicAdd(&dst[DS_DEC_CAND], (unsigned long long)c);
    atomicAdd(&dst[DS_DEC_ACC], (unsigned long long)nacc);
  }
}

__global__ __launch_bounds__(1024) void k_dec_small(MoveCtx m, CandOut cand, u32* minUto, u32* minUfrom,
                                                     u32* minAto, u32* minAfrom, u32* mark, u32* next,
                                                     u32* next_cnt, u32* status, unsigned long long* dst,
                                                     u32* cur_cnt, u32* bucket_counts, FlatMarks flat,
                                                     u32* round_cand) {
  __shared__ DecShared sh;
  dec_small_body(m, cand, minUto, minUfrom, minAto, minAfrom, mark, next, next_cnt, status, dst, cur_cnt,
                 bucket_counts, flat, round_cand, sh);
}

// Marks of long neighbour lists over the whole grid: positions p < *flat.c
// with flat.pre the exclusive prefix of their list lengths (flat.pre[c] =
// total); next-active vertices are staged per workgroup and appended with one
// device atomic per flush.
template <u32 BS>
struct MarksShared {
  u32 sv_[2 * kMarkPer * BS];  // two passes of next-active marks
  u32 scnt, sbase, span[2];
};
template <u32 BS>
__device__ __forceinline__ void marks_flat_body(const MoveCtx& m, const FlatMarks& flat, u32* mark, u32* next,
                                                u32* next_cnt, MarksShared<BS>& sh) {
  // kMarkPer entries per thread and pass, their gathers issued together so
  // that more of them are in flight; the stage holds two passes of marks
  constexpr u32 R = kMarkPer;
  constexpr u32 kStageN = 2 * R * BS;
  u32* sv_ = sh.sv_;
  u32& scnt = sh.scnt;
  u32& sbase = sh.sbase;
  u32* span = sh.span;
  const u32 c = *flat.c;
  if (c == 0) return;
  if (threadIdx.x == 0) scnt = 0;
  __syncthreads();
  const u64 total = flat.pre[c];
  const u64 per = u64(blockDim.x) * R;
  for (u64 base = u64(blockIdx.x) * per; base < total; base += u64(gridDim.x) * per) {
    // the segments of the pass's first and last entries bound every thread's
    // search, so the per-thread searches take a few cached steps
    if (threadIdx.x == 0 || threadIdx.x == blockDim.x - 1) {
      const u64 last = base + per - 1;
      const u64 e = threadIdx.x == 0 ? base : (last < total ? last : total - 1);
      u32 lo = 0, hi = c;
      while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (flat.pre[mid] <= e) lo = mid;
        else hi = mid;
      }
      span[threadIdx.x == 0 ? 0 : 1] = lo;
    }
    __syncthreads();
    u32 uu[R], vv[R], ct[R], fu[R], su[R], sv[R];
    u64 ee[R];
    u32 seg[R];
    bool ok[R], rep[R], need[R];
#pragma unroll
    for (u32 r = 0; r < R; ++r) {
      const u64 t = base + u64(r) * blockDim.x + threadIdx.x;
      ok[r] = t < total;
      uu[r] = vv[r] = ct[r] = seg[r] = 0;
      ee[r] = 0;
      rep[r] = false;
      if (ok[r]) {
        u32 lo = span[0], hi = span[1] + 1;
        while (hi - lo > 1) {
          const u32 mid = (lo + hi) >> 1;
          if (flat.pre[mid] <= t) lo = mid;
          else hi = mid;
        }
        seg[r] = lo;
        vv[r] = flat.node[lo];
        ct[r] = flat.to[lo];
        rep[r] = flat.kd && flat.kd[lo];
        ee[r] = m.g.row_off[vv[r]] + (t - flat.pre[lo]);
        uu[r] = ld_row(m.g.col + ee[r]);
      }
    }
    // a repeat move's K marks are done, and its next-active mark of a
    // neighbour that is already claimed needs no gather
#pragma unroll
    for (u32 r = 0; r < R; ++r) need[r] = ok[r] && !(rep[r] && marked(mark, uu[r]));
#pragma unroll
    for (u32 r = 0; r < R; ++r) {
      fu[r] = su[r] = sv[r] = 0;
      if (need[r]) {
        if (m.fs) {
          const u64 x = m.fs[uu[r]];
          fu[r] = u32(x);
          su[r] = u32(x >> 32);
        } else {
          fu[r] = m.f[uu[r]];
          su[r] = m.s[uu[r]];
        }
        if (!rep[r]) sv[r] = m.s[vv[r]];
      }
    }
#pragma unroll
    for (u32 r = 0; r < R; ++r) {
      if (!ok[r]) continue;
      const u32 u = uu[r], v = vv[r];
      if (m.ht.bits && flat.from && u != v) ht_move(m.ht, u, flat.from[seg[r]], ct[r], ld_row(m.g.w + ee[r]));
      if (need[r] && fu[r] != ct[r] && claim(mark, u)) sv_[atomicAdd(&scnt, 1u)] = u;
      if (need[r] && !rep[r] && su[r] == sv[r]) {
        m.K[u] = 1;
        m.K[v] = 1;
      }
    }
    __syncthreads();
    if (scnt > kStageN - R * BS) {
      if (threadIdx.x == 0) sbase = atomicAdd(next_cnt, scnt);
      __syncthreads();
      for (u32 i = threadIdx.x; i < scnt; i += blockDim.x) next[sbase + i] = sv_[i];
      __syncthreads();
      if (threadIdx.x == 0) scnt = 0;
      __syncthreads();
    }
  }
  __syncthreads();
  if (scnt) {
    if (threadIdx.x == 0) sbase = atomicAdd(next_cnt, scnt);
    __syncthreads();
    for (u32 i = threadIdx.x; i < scnt; i += blockDim.x) next[sbase + i] = sv_[i];
  }
}
template <u32 BS>
__global__ __launch_bounds__(BS) void k_marks_flat(MoveCtx m, FlatMarks flat, u32* mark, u32* next,
                                                   u32* next_cnt) {
  __shared__ MarksShared<BS> sh;
  marks_flat_body<BS>(m, flat, mark, next, next_cnt, sh);
}
// ------------------------------------------------------------------ persistent rounds
// A level's throughput rounds in one cooperative kernel of 1,024-thread
// workgroups (at most one per CU; a grid barrier costs about 0.13 us per
// workgroup, so the grid is small and the lists it takes are short).  Per
// round: the proposals (a wavefront per listed vertex, its neighbour
// communities in up to 512 LDS slots), a grid barrier, the decoupling with its
// apply and in-block marks by workgroup 0 (dec_small_body), a grid barrier,
// and, when the decoupling left long neighbour lists, the grid-wide marks and
// a grid barrier.  The rounds are the per-round launches' rounds (the same
// device bodies, with work counters of their own) without a host round trip
// between them.  A round whose keys do not fit the slots or whose candidates
// exceed the decoupling workgroup stops the kernel before anything of it is
// applied; the host runs that round with the per-round launches and may start
// the kernel again for a later round, as it does after a round that lists more
// than kPersistActive vertices.
constexpr u32 kPersistBS = 1024;
constexpr u32 kPersistG = 64, kPersistCap = 512;
constexpr u32 kPersistDeg = 256;       // rows of at most this many entries: tables of kPersistCap slots
constexpr u32 kPersistActive = 16384;  // longer lists take the per-round launches (wider grids)
struct PersistArgs {
  u32* list[2];  // the alternating frontier lists and their counts
  u32* lcnt[2];
  int ci0;     // the list of the first round
  u32 rounds;  // at most this many rounds
  u64 mark_words;
  u32* mark;
  CandOut cand;
  u32 *minUto, *minUfrom, *minAto, *minAfrom;
  u32* status;
  u32* bcounts;
  u32* ccand;
  u32* ovf;
  u32* ovf_cnt;
  FlatMarks fm;
  unsigned long long* dstat;
  // [0] rounds completed, [1] 1 when the rounds from [0] on are left to the
  // per-round launches, [2] vertices evaluated and [3] moves accepted in the
  // completed rounds, [4] the first round's list count
  unsigned long long* out;
};
union PersistShared {
  DecShared dec;
  TabShared<kPersistG, kPersistCap, kPersistBS> tab;
  MarksShared<kPersistBS> marks;
};
__global__ __launch_bounds__(kPersistBS) void k_move_rounds(MoveCtx m, MovePolicyT pol, PersistArgs a) {
  __shared__ PersistShared sh;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  unsigned long long* pst = a.dstat + DS_PR_PROP;
  unsigned long long* dst = a.dstat + (DS_PR_DEC - DS_DEC_CAND);  // dst[DS_DEC_CAND] is dstat[DS_PR_DEC]
  {
    // the first list, read by every workgroup before any barrier (nothing
    // changes it before the first one): an empty one ends the level, a long
    // one goes to the per-round launches
    const u32 c0 = *a.lcnt[a.ci0 & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      // out[] is this thread's alone: it starts from zero here
      a.out[0] = 0;
      a.out[1] = c0 > kPersistActive ? 1 : 0;
      a.out[2] = 0;
      a.out[3] = 0;
      a.out[4] = c0;
    }
    if (c0 == 0 || c0 > kPersistActive) return;
  }
  TabTally tally;
  for (u32 r = 0; r < a.rounds; ++r) {
    const int ci = (a.ci0 + int(r)) & 1;
    const u32* cur = a.list[ci];
    u32* nxt = a.list[ci ^ 1];
    u32* cur_cnt = a.lcnt[ci];
    u32* next_cnt = a.lcnt[ci ^ 1];
    // the claim bitmap of this round's marks, and the proposals
    for (u64 w = gtid(); w < a.mark_words; w += gstride()) a.mark[w] = 0;
    tab_body<MovePolicyT, kPersistG, kPersistCap, kPersistBS>(pol, m.g, cur, cur_cnt, a.ovf, a.ovf_cnt, a.cand, pst,
                                                              pst, sh.tab, nullptr, &tally);
    grid.sync();
    // every workgroup reads the same counts here (workgroup 0 clears the
    // candidate count only on the way on), so all stop or go on together
    if (*a.ovf_cnt != 0 || *a.cand.cnt > kDecSmall) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.out[0] = r;
        a.out[1] = 1;
      }
      break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.out[2] += *cur_cnt;  // read before the decoupling clears it
    if (blockIdx.x == 0)
      dec_small_body(m, a.cand, a.minUto, a.minUfrom, a.minAto, a.minAfrom, a.mark, nxt, next_cnt, a.status, dst,
                     cur_cnt, a.bcounts, a.fm, a.ccand, sh.dec);
    grid.sync();
    // the grid-wide marks only when the decoupling workgroup left rows to them
    // (the count it set stays until the next round's decoupling)
    if (*a.fm.c != 0) {
      marks_flat_body<kPersistBS>(m, a.fm, a.mark, nxt, next_cnt, sh.marks);
      grid.sync();
    }
    const u32 acc = a.status[1], nx = *next_cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.out[3] += acc;
      a.out[0] = r + 1;
      if (acc != 0 && nx > kPersistActive) a.out[1] = 1;
    }
    // throughput.rs:177-180, nothing left to evaluate, or a list for the
    // per-round launches' wider grids
    if (acc == 0 || nx == 0 || nx > kPersistActive) break;
  }
  const u32 idx[3] = {0, 1, 2};
  unsigned long long vals[3] = {tally.nv, tally.ne, tally.ncand};
  block_add_multi(pst, idx, vals);
}

// workgroups of the persistent rounds per device (0: not available)
static u32 g_persist_grid[64];
static bool g_persist_known[64];
static u32 persist_grid() {
  u32* grid = g_persist_grid;
  bool* known = g_persist_known;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (!known[dev]) {
    known[dev] = true;
    int coop = 0, cus = 0, per = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && coop &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&k_move_rounds),
                                                     int(kPersistBS), 0) == hipSuccess &&
        per > 0 && cus > 0)
      grid[dev] = u32(per) * u32(cus);
    (void)hipGetLastError();
  }
  return grid[dev];
}
// A cooperative launch the runtime refused: the device's rounds take the
// per-round launches from then on (the same results), and stderr says why, once.
static void persist_refused(hipError_t e, hipError_t before) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
  if (g_persist_grid[dev])
    std::fprintf(stderr, "[hl] device %d refused the cooperative rounds kernel (%s; earlier error state: %s); "
                 "the per-round launches run from here on\n", dev, hipGetErrorString(e), hipGetErrorString(before));
  g_persist_grid[dev] = 0;
}

// The moves' cache updates over the grid: entry t of the moved vertices'
// concatenated rows (one thread each) updates v's entry in neighbour u's row.
__global__ __launch_bounds__(256) void k_ck_flat(MoveCtx m, FlatMarks flat) {
  __shared__ u32 span[2];
  const u32 c = *flat.ck_c;
  if (c == 0) return;
  const u64 total = flat.pre[c];
  for (u64 base = u64(blockIdx.x) * blockDim.x; base < total; base += u64(gridDim.x) * blockDim.x) {
    const u64 t = base + threadIdx.x;
    if (threadIdx.x == 0 || threadIdx.x == blockDim.x - 1) {
      const u64 e = threadIdx.x == 0 ? base : (t < total ? t : total - 1);
      u32 lo = 0, hi = c;
      while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (flat.pre[mid] <= e) lo = mid;
        else hi = mid;
      }
      span[threadIdx.x == 0 ? 0 : 1] = lo;
    }
    __syncthreads();
    if (t < total) {
      u32 lo = span[0], hi = span[1] + 1;
      while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (flat.pre[mid] <= t) lo = mid;
        else hi = mid;
      }
      const u32 v = flat.node[lo];
      const u32 u = m.g.col[m.g.row_off[v] + (t - flat.pre[lo])];
      ck_update(m, u, v, flat.to[lo]);
    }
    __syncthreads();
  }
}
__global__ void k_acc_degrees(u64 c, const u32* node, const u32* state, const u32* row_len, u64* deg) {
  for (u64 p = gtid(); p < c; p += gstride()) deg[p] = state[p] == kAcc ? row_len[node[p]] : 0;
}

// zero the per-round counters the kernels above leave set
// the round's counts back to zero (null pointers skipped); also *set = val and
// zero2[0..1] = 0 where given
__global__ void k_round_reset(u32* cur_cnt, u32* ccnt, u32* bucket_counts, u32* set = nullptr, u32 val = 0,
                              u32* zero2 = nullptr) {
  if (threadIdx.x == 0) {
    if (cur_cnt) *cur_cnt = 0;
    if (ccnt) *ccnt = 0;
    clear_bucket_counts(bucket_counts);
    if (set) *set = val;
    if (zero2) zero2[0] = zero2[1] = 0;
  }
}

// ------------------------------------------------------------------ large candidate sets
__global__ void k_cand_node_keys(CandOut cand, u32 c, u32* keys, u32* vals) {
  for (u64 i = gtid(); i < c; i += gstride()) {
    keys[i] = cand.node[i];
    vals[i] = u32(i);
  }
}
__global__ void k_cand_gain_keys(const u32* order1, u32 c, CandOut cand, u64* keys) {
  for (u64 i = gtid(); i < c; i += gstride()) keys[i] = ~(u64)__double_as_longlong(cand.gain[order1[i]]);
}
__global__ void k_cand_arrange(const u32* order2, const u32* order1, u32 c, CandOut cand, const u32* f, u32* node,
                               u32* from, u32* to, u32* state, unsigned long long* dst) {
  if (gtid() == 0) atomicAdd(&dst[DS_MULTI_CAND], (unsigned long long)c);
  for (u64 p = gtid(); p < c; p += gstride()) {
    u32 i = order1[order2[p]];
    u32 v = cand.node[i];
    node[p] = v;
    from[p] = f[v];
    to[p] = cand.to[i];
    state[p] = kUndec;
  }
}
// The priority-round kernels take the previous round's undecided count
// (prev, or null for the first round of a chunk) and do nothing once it is
// zero, so a chunk of rounds is enqueued with one read-back at its end.
__global__ void k_reset4(u64 c, const u32* from, const u32* to, u32* a, u32* b, u32* x, u32* y,
                         const u32* prev) {
  if (prev && *prev == 0) return;
  for (u64 p = gtid(); p < c; p += gstride()) {
    u32 f = from[p], t = to[p];
    a[f] = kNone; a[t] = kNone;
    b[f] = kNone; b[t] = kNone;
    x[f] = kNone; x[t] = kNone;
    y[f] = kNone; y[t] = kNone;
  }
}
// The minima only fall during a kernel, so a plain read at or below x means
// the atomic would change nothing.
__device__ __forceinline__ void min_if_lower(u32* a, u32 x) {
  if (*a > x) atomicMin(a, x);
}
// Per workgroup pass of kMarkChunk candidates, the minima are first taken in
// LDS tables (one per array, keyed by community), so a community that many
// candidates share (a giant one) takes one device atomic per pass instead of
// one per candidate.
constexpr u32 kMarkChunk = 512;
constexpr u32 kMarkSlots = 1024;  // >= 2 kMarkChunk: at most half full
__device__ __forceinline__ void lds_min_insert(u32* K, u32* M, u32 key, u32 p) {
  u32 h = hash32(key) & (kMarkSlots - 1);
  while (true) {
    const u32 old = atomicCAS(&K[h], kNone, key);
    if (old == kNone || old == key) {
      atomicMin(&M[h], p);
      return;
    }
    h = (h + 1) & (kMarkSlots - 1);
  }
}
__global__ __launch_bounds__(256) void k_mark4(u64 c, const u32* from, const u32* to, const u32* state, u32* minUto,
                                               u32* minUfrom, u32* minAto, u32* minAfrom, const u32* prev) {
  __shared__ u32 K[4 * kMarkSlots], M[4 * kMarkSlots];
  if (prev && *prev == 0) return;
  for (u64 base = u64(blockIdx.x) * kMarkChunk; base < c; base += u64(gridDim.x) * kMarkChunk) {
    for (u32 q = threadIdx.x; q < 4 * kMarkSlots; q += blockDim.x) {
      K[q] = kNone;
      M[q] = kNone;
    }
    __syncthreads();
    for (u32 r = threadIdx.x; r < kMarkChunk; r += blockDim.x) {
      const u64 p = base + r;
      if (p >= c) break;
      const u32 st = state[p];
      if (st == kUndec) {
        lds_min_insert(K, M, to[p], u32(p));
        lds_min_insert(K + kMarkSlots, M + kMarkSlots, from[p], u32(p));
      } else if (st == kAcc) {
        lds_min_insert(K + 2 * kMarkSlots, M + 2 * kMarkSlots, to[p], u32(p));
        lds_min_insert(K + 3 * kMarkSlots, M + 3 * kMarkSlots, from[p], u32(p));
      }
    }
    __syncthreads();
    for (u32 q = threadIdx.x; q < 4 * kMarkSlots; q += blockDim.x) {
      const u32 key = K[q];
      if (key == kNone) continue;
      const u32 a = q / kMarkSlots;
      u32* dst = a == 0 ? minUto : a == 1 ? minUfrom : a == 2 ? minAto : minAfrom;
      min_if_lower(&dst[key], M[q]);
    }
    __syncthreads();
  }
}
__global__ void k_decide4(u64 c, const u32* from, const u32* to, u32* state, const u32* minUto, const u32* minUfrom,
                          const u32* minAto, const u32* minAfrom, u32* undec, const u32* prev) {
  if (prev && *prev == 0) return;
  u32 local = 0;
  for (u64 p = gtid(); p < c; p += gstride()) {
    if (state[p] != kUndec) continue;
    u32 f = from[p], t = to[p], pp = u32(p);
    if (minAto[f] < pp || minAfrom[t] < pp)
      state[p] = kRej;
    else if (minUto[f] < pp || minUfrom[t] < pp)
      ++local;
    else
      state[p] = kAcc;
  }
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(undec, local);
}
// A chain of conflicts longer than the priority rounds allow: the undecided
// rest is settled in priority order by one thread.  An accepted candidate
// never has an undecided conflicting candidate before it (it would have
// waited), so the flags of every accepted candidate, later ones included,
// touch an undecided candidate's checks only through earlier ones, and the
// result is the sequential greedy's (throughput.rs:204-236).  The flag arrays
// hold kNone at every candidate's communities on entry (k_reset4); a flag is
// any other value.  undec: 1 for each undecided candidate, for the ordered
// select.
__global__ void k_settle_prep(u64 c, const u32* from, const u32* to, const u32* state, u32* emit, u32* recv,
                              u8* undec) {
  for (u64 p = gtid(); p < c; p += gstride()) {
    const u32 st = state[p];
    undec[p] = st == kUndec;
    if (st == kAcc) {
      emit[from[p]] = 0;
      recv[to[p]] = 0;
    }
  }
}
__global__ void k_settle_seq(const u32* list, const unsigned long long* cnt, const u32* from, const u32* to,
                             u32* state, u32* emit, u32* recv, unsigned long long* dst) {
  if (gtid() != 0) return;
  const u64 n = *cnt;
  dst[DS_SETTLED] += n;
  for (u64 i = 0; i < n; ++i) {
    const u32 p = list[i];
    const u32 f = from[p], t = to[p];
    if (recv[f] != kNone || emit[t] != kNone) {
      state[p] = kRej;
    } else {
      state[p] = kAcc;
      emit[f] = 0;
      recv[t] = 0;
    }
  }
}
__global__ void k_apply4(u64 c, const u32* node, const u32* from, const u32* to, const u32* state, MoveCtx m,
                         u32* acc_list, u32* acc_cnt, unsigned long long* dst, u8* kd) {
  unsigned long long na = 0, nd = 0;
  for (u64 base = u64(blockIdx.x) * blockDim.x; base < c; base += gstride()) {
    u64 p = base + threadIdx.x;
    bool acc = p < c && state[p] == kAcc;
    if (p < c) kd[p] = acc ? m.B[node[p]] : u8(0);  // moved in an earlier round of this level
    if (acc) {
      u32 v = node[p];
      ++na;
      nd += m.g.row_len[v];
      double kv = m.k[v];
      m.f[v] = to[p];
      if (m.fs) *fs_f(m.fs, v) = to[p];
      atomicAdd(&m.tot[from[p]], -kv);
      atomicAdd(&m.tot[to[p]], kv);
      m.B[v] = 1;
    }
    u32 q = wave_append(acc, acc_cnt);
    if (acc) acc_list[q] = u32(p);
  }
  block_add_u64(&dst[DS_MULTI_ACC], na);
  block_add_u64(&dst[DS_MULTI_DEG], nd);
}

// ------------------------------------------------------------------ refinement helpers
__global__ void k_ref_targets(CandOut cand, u8* target) {
  const u32 c = *cand.cnt;
  for (u64 i = gtid(); i < c; i += gstride()) target[cand.to[i]] = 1;
}
__global__ void k_ref_apply(CandOut cand, const u8* target, u32* s, const double* k, SubRef sub) {
  const u32 c = *cand.cnt;
  for (u64 i = gtid(); i < c; i += gstride()) {
    const u32 v = cand.node[i], from = s[v], to = cand.to[i];
    if (target && target[from]) continue;  // a target singleton keeps its place
    s[v] = to;
    if (sub.size) {
      atomicSub(&sub.size[from], 1u);
      atomicAdd(&sub.size[to], 1u);
      atomicAdd(&sub.deg[from], -k[v]);
      atomicAdd(&sub.deg[to], k[v]);
    }
  }
}
__global__ void k_ref_clear_targets(CandOut cand, u8* target) {
  const u32 c = *cand.cnt;
  for (u64 i = gtid(); i < c; i += gstride()) target[cand.to[i]] = 0;
}

}  // namespace

// ====================================================================== host side

static unsigned fixed_grid(u64 work, unsigned bs = 256, unsigned cap = 2048) {
  u64 b = (work + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return unsigned(b);
}

void MoveScratch::ensure(u64 n, u64 comm_cap, hipStream_t s) {
  if (!dstat.p) {
    dstat.reserve(DS_COUNT, s);
    HL_CHECK(hipMemsetAsync(dstat.p, 0, DS_COUNT * sizeof(unsigned long long), s));
  }
  mark.reserve(n / 32 + 2, s);
  mark_n = u32(n);
  list_a.reserve(n + 1, s);
  list_b.reserve(n + 1, s);
  blists.reserve((kHtList + 1) * (n + 1), s);
  cnode.reserve(n + 1, s);
  cto.reserve(n + 1, s);
  cgain.reserve(n + 1, s);
  counters.reserve(32, s);
  mnode.reserve(kDecSmall + 1, s);
  mto.reserve(kDecSmall + 1, s);
  mfrom.reserve(kDecSmall + 1, s);
  mpre.reserve(kDecSmall + 2, s);
  mkd.reserve(kDecSmall + 1, s);
  if (!snap) {
    HL_CHECK(hipHostMalloc((void**)&snap, 2 * 16 * sizeof(u32), hipHostMallocDefault));
    for (hipEvent_t& e : snap_ev) HL_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (minUto.cap < comm_cap) {
    minUto.reserve(comm_cap, s);
    minUfrom.reserve(comm_cap, s);
    minAto.reserve(comm_cap, s);
    minAfrom.reserve(comm_cap, s);
  }
}

MoveScratch::~MoveScratch() {
  if (snap) (void)hipHostFree(snap);
  for (hipEvent_t e : snap_ev)
    if (e) (void)hipEventDestroy(e);
}

void MoveScratch::trim() {
  for (DevBuf<u32>* b : {&tk, &hovf, &ck, &k32, &k32b, &v32, &o1, &o2, &node, &from, &to, &state, &acc})
    b->release();
  ht_slot.release();
  tv.release();
  k64.release();
  k64b.release();
}

void MoveScratch::fresh_marks(hipStream_t s) { fill_u32(mark.p, 0, u64(mark_n) / 32 + 1, s); }

template <class P>
static void propose_round(const P& pol, const Csr& g, const u32* list, const u32* cnt_dev, u64 list_cap, u32 maxdeg,
                          MoveScratch& ms, const Thresholds& th, unsigned long long* st, unsigned long long* tst,
                          hipStream_t s, u64 grid_work = 0) {
  // list_cap: at least the list's count, the stride of the bucket lists;
  // grid_work: the count the grids are sized for (every kernel strides over
  // the count it reads on the device, so a short estimate only costs time)
  if (grid_work == 0) grid_work = list_cap;
  // counters [4..4 + kBuckets) bucket counts, [kOvfSlot..kOvfSlot + kOvfLists)
  // overflow counts; all zero on entry.  st: the proposal block of the device
  // work counters; tst: the first-try block (3 counters per bucket)
  u32* counts = ms.counters.p + 4;
  const u32 small = th.small_deg > SMALL_MAX2 ? SMALL_MAX2 : th.small_deg;
  const u32 wave = th.wave_deg > WAVE_HW2 / 2 ? WAVE_HW2 / 2 : th.wave_deg;
  const u32 half = std::max(small, std::min(kHalfDeg, wave));
  CandOut out{ms.cnode.p, ms.cto.p, ms.cgain.p, ms.counters.p + 1};
  unsigned long long* cst = st + 2 * kBuckets;
  {
    HL_PROF(P_PROP_SMALL, 0, s);
    hipLaunchKernelGGL(k_prop2_first<P>, fixed_grid(grid_work), 256, 0, s, pol, g, list, cnt_dev, small, half, wave,
                       ms.blists.p, list_cap, counts, out, st);
    HL_LAUNCH_CHECK();
  }
  if constexpr (P::kTables) {
    if (pol.ht.bits) {
      HL_PROF(P_PROP_TABLE, 0, s);
      hipLaunchKernelGGL(k_prop2_ht<P>, dim3(fixed_grid(grid_work, 4, 2048)), dim3(256), 0, s, pol, g,
                         ms.blists.p + kHtList * list_cap, counts + kHtList, out, ms.dstat.p + DS_HT);
      HL_LAUNCH_CHECK();
    }
  }
  // each bucket past the first: small tables first, then the tables sized for
  // the degree for the vertices listed as overflowing them
  u32* ovf_cnt = ms.counters.p + kOvfSlot;
  if (maxdeg > small) ms.hovf.reserve(kOvfLists * list_cap, s);
  u32* ovf[kOvfLists];
  for (u32 i = 0; i < kOvfLists; ++i) ovf[i] = ms.hovf.p + i * list_cap;
  if (maxdeg > small) {
    {
      HL_PROF(P_TRY_HALF, 0, s);
      hipLaunchKernelGGL((k_prop2_tab<P, 32, 64>), dim3(fixed_grid(grid_work, 8, kTabBlocks)), dim3(256), 0, s, pol,
                         g, ms.blists.p + list_cap, counts + 1, ovf[0], ovf_cnt + 0, out, cst + 1, tst + 3 * 1);
      HL_LAUNCH_CHECK();
    }
    {
      HL_PROF(P_PROP_WAVE, 0, s);
      hipLaunchKernelGGL((k_prop2_wave<P, 32, 2 * kHalfDeg>), dim3(fixed_grid(grid_work, 8, kHalfBlocks)), dim3(256),
                         0, s, pol, g, ovf[0], ovf_cnt + 0, out, cst + 1);
      HL_LAUNCH_CHECK();
    }
    if (maxdeg > half) {
      {
        HL_PROF(P_TRY_WAVE, 0, s);
        hipLaunchKernelGGL((k_prop2_tab<P, 64, 128>), dim3(fixed_grid(grid_work, 4, kTabBlocks)), dim3(256), 0, s,
                           pol, g, ms.blists.p + 2 * list_cap, counts + 2, ovf[1], ovf_cnt + 1, out, cst + 2,
                           tst + 3 * 2);
        HL_LAUNCH_CHECK();
      }
      HL_PROF(P_PROP_WAVE, 0, s);
      hipLaunchKernelGGL((k_prop2_wave<P, 64, WAVE_HW2>), dim3(fixed_grid(grid_work, 4, kWaveBlocks)), dim3(256), 0, s,
                         pol, g, ovf[1], ovf_cnt + 1, out, cst + 2);
      HL_LAUNCH_CHECK();
    }
  }
  if (maxdeg > wave) {
    {
      HL_PROF(P_TRY_LDS, 0, s);
      hipLaunchKernelGGL((k_prop2_hub<P, 256, 2048>), dim3(1536), dim3(256), 0, s, pol, g,
                         ms.blists.p + 3 * list_cap, counts + 3, ovf[2], ovf_cnt + 2, out, cst + 3, tst + 3 * 3);
      HL_LAUNCH_CHECK();
    }
    HL_PROF(P_PROP_LDS, 0, s);
    hipLaunchKernelGGL(k_prop2_lds<P>, dim3(fixed_grid(grid_work, 1, 256)), dim3(1024), 0, s, pol, g, ovf[2],
                       ovf_cnt + 2, out, cst + 3);
    HL_LAUNCH_CHECK();
  }
  if (maxdeg > kLdsDeg) {
    // hubs: LDS tables first, then HBM tables (one per resident workgroup, at
    // most kBigTableBytes in all) for the vertices whose keys overflow them
    {
      HL_PROF(P_TRY_HUB, 0, s);
      hipLaunchKernelGGL((k_prop2_hub<P, 512, 4096>), dim3(768), dim3(512), 0, s, pol, g,
                         ms.blists.p + 4 * list_cap, counts + 4, ovf[3], ovf_cnt + 3, out, cst + 4, tst + 3 * 4);
      HL_LAUNCH_CHECK();
    }
    if (maxdeg > kBigDeg) {
      HL_PROF(P_TRY_GIANT, 0, s);
      hipLaunchKernelGGL((k_prop2_hub<P, 1024, 8192>), dim3(256), dim3(1024), 0, s, pol, g,
                         ms.blists.p + 5 * list_cap, counts + 5, ovf[3], ovf_cnt + 3, out, cst + 5, tst + 3 * 5);
      HL_LAUNCH_CHECK();
    }
    const u32 tcap = big_cap(maxdeg);
    const u64 per = u64(tcap) * (sizeof(u32) + sizeof(double));
    const unsigned nb = unsigned(std::max<u64>(1, std::min<u64>(kOvfBlocks, kBigTableBytes / per)));
    ms.tk.reserve(u64(nb) * tcap, s);
    ms.tv.reserve(u64(nb) * tcap, s);
    HL_PROF(P_PROP_BIG, 0, s);
    hipLaunchKernelGGL(k_prop2_big<P>, dim3(nb), dim3(1024), 0, s, pol, g, ovf[3], ovf_cnt + 3, ms.tk.p, ms.tv.p,
                       tcap, out, cst + 4);
    HL_LAUNCH_CHECK();
  }
}

// the grid-wide neighbour marks at the workgroup shape the thresholds name
static void launch_marks(const Thresholds& th, const MoveCtx& m, const FlatMarks& fm, u32* mark, u32* nxt,
                         u32* next_cnt, hipStream_t s) {
  const unsigned grid = th.marks_grid ? th.marks_grid : kMarksBlocks;
  if (th.marks_block == 128)
    hipLaunchKernelGGL(k_marks_flat<128>, grid, 128, 0, s, m, fm, mark, nxt, next_cnt);
  else if (th.marks_block == 512)
    hipLaunchKernelGGL(k_marks_flat<512>, grid, 512, 0, s, m, fm, mark, nxt, next_cnt);
  else
    hipLaunchKernelGGL(k_marks_flat<256>, grid, 256, 0, s, m, fm, mark, nxt, next_cnt);
  HL_LAUNCH_CHECK();
}

u32 movement_throughput(const MoveCtx& m_in, const Rec* d, u64 k, u32 maxdeg, int max_rounds, MoveScratch& ms,
                        const Thresholds& th, Temp& tmp, MoveStats* st, hipStream_t s, const RoundHook* hook) {
  MoveCtx m = m_in;
  const u32 n = m.g.n;
  ms.ensure(n, m.comm_cap, s);
  ms.fs.reserve(u64(n) + 1, s);
  hipLaunchKernelGGL(k_pack_fs, fixed_grid(n, 256, 4096), 256, 0, s, m.f, m.s, n, ms.fs.p);
  HL_LAUNCH_CHECK();
  m.fs = ms.fs.p;
  if (m.pool) {
    ms.ck.reserve(m.pool, s);
    m.ck = ms.ck.p;
    hipLaunchKernelGGL(k_key_cache, fixed_grid(n, 256, 4096), 256, 0, s, m.g, m.f, m.ck);
    HL_LAUNCH_CHECK();
  }
  // counters: [0], [2] list counts (alternating), [1] candidates, [4..10) bucket
  // counts, [13] flat mark count, [14] overflow flag, [15] accepted moves,
  // [16..20) table overflow counts; a round leaves all but the next list's
  // count at zero
  u32* lcnt[2] = {ms.counters.p + 0, ms.counters.p + 2};
  u32* ccnt = ms.counters.p + 1;
  u32* bcounts = ms.counters.p + 4;
  u32* status = ms.counters.p + 14;
  int ci = 0;
  u32* cnt_a = lcnt[0];
  u32* cur = ms.list_a.p;
  u32* nxt = ms.list_b.p;
  if (k == 0 && !m.gid) {
    fill_u32(ms.counters.p, 0, 32, s);
    hipLaunchKernelGGL(k_iota_list, fixed_grid(n, 256, 4096), 256, 0, s, cur, n, cnt_a);
    HL_LAUNCH_CHECK();
  } else {
    // the counters and the claim bitmap cleared in one launch
    const u64 words = u64(ms.mark_n) / 32 + 1;
    hipLaunchKernelGGL(k_move_clear, fixed_grid(words, 256, 4096), 256, 0, s, ms.counters.p, ms.mark.p, words);
    HL_LAUNCH_CHECK();
    if (k) {
      HL_PROF(P_SEED, 24 * k, s);
      hipLaunchKernelGGL(k_seed2, fixed_grid(k), 256, 0, s, d, k, m.f, m.s, m.A, m.K, ms.mark.p, cur, cnt_a);
      HL_LAUNCH_CHECK();
    }
    if (m.gid) {
      hipLaunchKernelGGL(k_activate_graphs, fixed_grid(n, 256, 4096), 256, 0, s, m.gid, n, m.part, m.gcnt, m.A,
                         ms.mark.p, cur, cnt_a);
      HL_LAUNCH_CHECK();
    }
  }
  if (m.ht.pool && !m.ck && maxdeg > std::min(th.wave_deg, WAVE_HW2 / 2)) {
    // hub tables: a cleared bitmap and an empty pool for this level's rounds
    const u64 words = (u64(n) + 32) / 32;
    ms.ht_bits.reserve(words, s);
    fill_u32(ms.ht_bits.p, 0, words, s);
    ms.ht_loc.reserve(n + 1, s);
    ms.ht_slot.reserve(m.ht.pool, s);
    ms.ht_used.reserve(1, s);
    hipLaunchKernelGGL(k_fill_u64_one, 1, 64, 0, s, ms.ht_used.p);
    HL_LAUNCH_CHECK();
    m.ht = HubTables{ms.ht_bits.p, ms.ht_loc.p, ms.ht_slot.p, ms.ht_used.p, m.ht.pool};
  } else {
    m.ht = HubTables{};
  }
  MovePolicyT pol{m.f, m.tot, m.gamma, n, m.ck, m.ht};
  u32 host_cnt = 0;
  u32 rounds = 0;
  // One round: proposals over `lst` (count lcnt[cidx], read on the device),
  // the one-workgroup decoupling with its marks into `out` (count
  // lcnt[cidx ^ 1]), then the grid-wide marks of long neighbour lists.  The
  // bucket lists have stride n + 1, so a round can be enqueued before its
  // count is known; grid_work sizes its grids.
  const u64 list_cap = u64(n) + 1;
  u32* ccand = ms.counters.p + 3;  // candidates of the round (k_dec_small)
  auto enqueue_round = [&](const u32* lst, u32* out, int cidx, u64 grid_work) {
    u32* c_cnt = lcnt[cidx];
    u32* n_cnt = lcnt[cidx ^ 1];
    propose_round(pol, m.g, lst, c_cnt, list_cap, maxdeg, ms, th, ms.dstat.p + DS_MOVE, ms.dstat.p + DS_TRY_MOVE, s,
                  grid_work);
    ms.fresh_marks(s);
    CandOut cand{ms.cnode.p, ms.cto.p, ms.cgain.p, ccnt};
    const FlatMarks fm{ms.mnode.p, ms.mto.p, ms.mpre.p, ms.counters.p + 13, ms.counters.p + 12, ms.mfrom.p, ms.mkd.p};
    {
      HL_PROF(P_DECOUPLE, 0, s);
      hipLaunchKernelGGL(k_dec_small, 1, 1024, 0, s, m, cand, ms.minUto.p, ms.minUfrom.p, ms.minAto.p,
                         ms.minAfrom.p, ms.mark.p, out, n_cnt, status, ms.dstat.p, c_cnt, bcounts, fm, ccand);
      HL_LAUNCH_CHECK();
    }
    if (m.ck) {
      HL_PROF(P_CACHE_UPDATE, 0, s);
      hipLaunchKernelGGL(k_ck_flat, kMarksBlocks, 256, 0, s, m, fm);
      HL_LAUNCH_CHECK();
    }
    {
      // neighbour lists too long for the decoupling workgroup: marks over the
      // grid (the kernel returns at once when the workgroup marked them all or
      // the candidates took the multi-kernel path: the count it reads is 0)
      HL_PROF(P_MARK_MOVED, 0, s);
      launch_marks(th, m, fm, ms.mark.p, out, n_cnt, s);
      HL_LAUNCH_CHECK();
    }
  };
  // The next round is enqueued before the host waits for this one's counters
  // when the last round's candidates fitted the one-workgroup decoupling with
  // room to spare (no served-form hook reads each round).  Such a round reads
  // this round's output count on the device: after the last round, or after
  // one that takes the multi-kernel path (whose k_dec_small leaves the list
  // and bucket counts at zero), it finds an empty list and does nothing, and
  // then runs again for real.
  const bool may_queue = !(hook && hook->fn);
  bool queued = false;
  u32 last_cand = kNone;
  int slot = 0;
  // Rounds in one cooperative kernel when they fit it: no hub tables or
  // neighbour cache at this level, rows of at most kPersistDeg entries, no
  // served-form hook reading every round, a list of at most kPersistActive
  // vertices.  The kernel runs from round `rounds` on until the level ends or
  // a round does not fit it; a round it left runs with the launches, and the
  // kernel is tried again once a later round's list fits.  While it may be
  // used, no round is enqueued ahead (the kernel starts from a known count).
  const u32 pgrid = th.persist ? std::min(persist_grid(), th.persist_blocks) : 0;
  const bool may_persist = pgrid && !m.ck && !m.ht.bits && !(hook && hook->fn) && maxdeg <= kPersistDeg;
  u32 left_round = kNone;  // the round the kernel last left to the launches
  bool refused = false;    // the runtime refused the cooperative launch
  u32 first_cnt = 0;       // the list count the kernel's first round read
  // false when the kernel could not be launched (nothing was enqueued)
  auto run_persist = [&]() -> bool {
    ms.hovf.reserve(u64(n) + 1, s);
    ms.pout.reserve(8, s);  // the kernel clears it
    PersistArgs a{};
    a.list[0] = ms.list_a.p;
    a.list[1] = ms.list_b.p;
    a.lcnt[0] = lcnt[0];
    a.lcnt[1] = lcnt[1];
    a.ci0 = ci;
    a.rounds = u32(max_rounds) - rounds;
    a.mark_words = u64(n) / 32 + 1;
    a.mark = ms.mark.p;
    a.cand = CandOut{ms.cnode.p, ms.cto.p, ms.cgain.p, ccnt};
    a.minUto = ms.minUto.p;
    a.minUfrom = ms.minUfrom.p;
    a.minAto = ms.minAto.p;
    a.minAfrom = ms.minAfrom.p;
    a.status = status;
    a.bcounts = bcounts;
    a.ccand = ccand;
    a.ovf = ms.hovf.p;
    a.ovf_cnt = ms.counters.p + kOvfSlot;  // zero: every round leaves it so
    a.fm = FlatMarks{ms.mnode.p, ms.mto.p, ms.mpre.p, ms.counters.p + 13, ms.counters.p + 12, ms.mfrom.p, ms.mkd.p};
    a.dstat = ms.dstat.p;
    a.out = ms.pout.p;
    void* args[] = {&m, &pol, &a};
    {
      HL_PROF(P_MOVE_ROUNDS, 0, s);
      const hipError_t before = hipGetLastError();
      const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_move_rounds), dim3(pgrid),
                                                      dim3(kPersistBS), args, 0, s);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        persist_refused(e, before);
        refused = true;
        return false;
      }
    }
    unsigned long long o[5] = {0, 0, 0, 0, 0};
    u32 cnts[3] = {0, 0, 0};  // counters [0] and [2]: the two lists' counts
    d2h_group({{o, ms.pout.p, sizeof(o)}, {cnts, lcnt[0], sizeof(cnts)}}, s);
    stream_wait_reads(s);
    const u32 done = u32(o[0]);
    first_cnt = u32(o[4]);
    rounds += done;
    if (st) {
      st->evaluated += o[2];
      st->moved_rounds += o[3];
      st->persist_rounds += done;
    }
    if (done & 1u) {
      std::swap(cur, nxt);
      ci ^= 1;
    }
    if (o[1]) {
      // the rounds from here on run with the launches; a round the kernel
      // left (its list and count untouched) runs again, without the
      // candidates and spills of the kernel's attempt
      host_cnt = cnts[2 * ci];
      fill_u32(ccnt, 0, 1, s);
      fill_u32(ms.counters.p + kOvfSlot, 0, kOvfLists, s);
      left_round = rounds;
    } else {
      host_cnt = 0;  // the last round accepted nothing, left nothing to evaluate, or was the last allowed
    }
    return true;
  };
  // The level's first list count: read by the kernel itself when it may take
  // the first rounds and the list is known to fit it (the seeding lists the
  // records' endpoints: at most 2k vertices, or all n when k = 0), so the host
  // does not wait before them; else read back.
  bool counted = false;
  if (may_persist && !m.gid && (k == 0 ? u64(n) : 2 * k) <= kPersistActive) {
    host_cnt = 1;  // not known here: the kernel reads it
    if (run_persist()) {
      counted = true;
      if (st) st->active = first_cnt;
    }
  }
  if (!counted) {
    d2h_async(&host_cnt, cnt_a, sizeof(u32), s);
    stream_wait_reads(s);
    if (st) st->active = host_cnt;
    if (hook && hook->fn) hook->fn(hook->ctx, cur, host_cnt);
  }
  while (host_cnt > 0 && int(rounds) < max_rounds) {
    if (may_persist && !refused && !queued && host_cnt <= kPersistActive && rounds != left_round && run_persist())
      continue;
    if (st) st->evaluated += host_cnt;
    u32* cur_cnt = lcnt[ci];
    u32* next_cnt = lcnt[ci ^ 1];
    CandOut cand{ms.cnode.p, ms.cto.p, ms.cgain.p, ccnt};
    if (!queued) enqueue_round(cur, nxt, ci, host_cnt);
    u32* snap = ms.snap + 16 * slot;
    HL_CHECK(hipMemcpyAsync(snap, ms.counters.p, 16 * sizeof(u32), hipMemcpyDeviceToHost, s));
    HL_CHECK(hipEventRecord(ms.snap_ev[slot], s));
    queued = may_queue && !may_persist && int(rounds) + 1 < max_rounds && last_cand <= kDecSmall / 2;
    if (queued) enqueue_round(nxt, cur, ci ^ 1, host_cnt);
    HL_CHECK(hipEventSynchronize(ms.snap_ev[slot]));
    g_alloc.syncs += 1;
    u32 h[16];
    std::memcpy(h, snap, sizeof(h));
    slot ^= 1;
    u32 accepted = h[15];
    u32 next_host = h[2 * (ci ^ 1)];
    last_cand = h[3];
    if (h[14] == 1) {
      queued = false;  // a round enqueued behind this one did nothing: it runs again
      // many candidates: sort by node, then stably by gain; priority rounds over the grid
      const u32 c = h[1];
      ms.k32.reserve(c, s);
      ms.v32.reserve(c, s);
      ms.o1.reserve(c, s);
      ms.k32b.reserve(c, s);
      ms.k64.reserve(c, s);
      ms.k64b.reserve(c, s);
      ms.o2.reserve(c, s);
      ms.node.reserve(c, s);
      ms.from.reserve(c, s);
      ms.to.reserve(c, s);
      ms.state.reserve(c, s);
      ms.acc.reserve(c, s);
      ms.kd.reserve(c, s);
      hipLaunchKernelGGL(k_cand_node_keys, fixed_grid(c), 256, 0, s, cand, c, ms.k32.p, ms.v32.p);
      HL_LAUNCH_CHECK();
      sort_pairs_u32(ms.k32.p, ms.k32b.p, ms.v32.p, ms.o1.p, c, 32, tmp, s);
      hipLaunchKernelGGL(k_cand_gain_keys, fixed_grid(c), 256, 0, s, ms.o1.p, c, cand, ms.k64.p);
      HL_LAUNCH_CHECK();
      iota_u32(ms.v32.p, 0, c, s);
      sort_pairs_u64(ms.k64.p, ms.k64b.p, ms.v32.p, ms.o2.p, c, 64, tmp, s);
      hipLaunchKernelGGL(k_cand_arrange, fixed_grid(c), 256, 0, s, ms.o2.p, ms.o1.p, c, cand, m.f, ms.node.p,
                         ms.from.p, ms.to.p, ms.state.p, ms.dstat.p);
      HL_LAUNCH_CHECK();
      // priority rounds in chunks of up to kDecChunk, each chunk read back once
      // (a round after one that decided everything does nothing)
      constexpr int kDecChunk = 4;
      ms.und.reserve(kDecChunk, s);
      int iters = 0;
      while (true) {
        const int chunk = std::min<int>(kDecChunk, int(th.dec_rounds) - iters);
        u32 und_h[kDecChunk] = {};
        {
          HL_PROF(P_DEC_MULTI, 0, s);
          fill_u32(ms.und.p, 0, u64(chunk), s);
          for (int r = 0; r < chunk; ++r) {
            const u32* prev = r ? ms.und.p + r - 1 : nullptr;
            hipLaunchKernelGGL(k_reset4, fixed_grid(c), 256, 0, s, u64(c), ms.from.p, ms.to.p, ms.minUto.p,
                               ms.minUfrom.p, ms.minAto.p, ms.minAfrom.p, prev);
            hipLaunchKernelGGL(k_mark4, fixed_grid(c), 256, 0, s, u64(c), ms.from.p, ms.to.p, ms.state.p,
                               ms.minUto.p, ms.minUfrom.p, ms.minAto.p, ms.minAfrom.p, prev);
            hipLaunchKernelGGL(k_decide4, fixed_grid(c), 256, 0, s, u64(c), ms.from.p, ms.to.p, ms.state.p,
                               ms.minUto.p, ms.minUfrom.p, ms.minAto.p, ms.minAfrom.p, ms.und.p + r, prev);
            HL_LAUNCH_CHECK();
          }
          d2h_async(und_h, ms.und.p, size_t(chunk) * sizeof(u32), s);
        }
        stream_wait_reads(s);
        int ran = chunk;  // the rounds that decided something: up to the first with none undecided
        for (int r = 0; r < chunk; ++r)
          if (und_h[r] == 0) {
            ran = r + 1;
            break;
          }
        if (st) st->decide_rounds += u32(ran);
        iters += ran;
        if (und_h[ran - 1] == 0) break;
        if (iters >= int(th.dec_rounds)) {
          // the undecided rest in priority order, against flags of the accepted
          // moves in minAfrom (sources) and minAto (targets); the sort buffers
          // are free by now: k32b holds the undecided flags, v32 their list,
          // k64b its count
          hipLaunchKernelGGL(k_reset4, fixed_grid(c), 256, 0, s, u64(c), ms.from.p, ms.to.p, ms.minUto.p,
                             ms.minUfrom.p, ms.minAto.p, ms.minAfrom.p, static_cast<const u32*>(nullptr));
          u8* undec = reinterpret_cast<u8*>(ms.k32b.p);
          hipLaunchKernelGGL(k_settle_prep, fixed_grid(c), 256, 0, s, u64(c), ms.from.p, ms.to.p, ms.state.p,
                             ms.minAfrom.p, ms.minAto.p, undec);
          HL_LAUNCH_CHECK();
          select_flagged_async(undec, c, ms.v32.p, tmp, ms.k64b.p, s);
          hipLaunchKernelGGL(k_settle_seq, 1, 64, 0, s, ms.v32.p,
                             reinterpret_cast<const unsigned long long*>(ms.k64b.p), ms.from.p, ms.to.p, ms.state.p,
                             ms.minAfrom.p, ms.minAto.p, ms.dstat.p);
          HL_LAUNCH_CHECK();
          break;
        }
      }
      u32* acc_cnt = status + 1;
      HL_CHECK(hipMemsetAsync(acc_cnt, 0, sizeof(u32), s));
      {
        HL_PROF(P_APPLY_MOVES, 0, s);
        hipLaunchKernelGGL(k_apply4, fixed_grid(c), 256, 0, s, u64(c), ms.node.p, ms.from.p, ms.to.p, ms.state.p, m,
                           ms.acc.p, acc_cnt, ms.dstat.p, ms.kd.p);
        HL_LAUNCH_CHECK();
      }
      {
        ms.mdeg.reserve(c, s);
        ms.mpre.reserve(u64(c) + 2, s);
        const FlatMarks fmm{ms.node.p, ms.to.p, ms.mpre.p, ms.counters.p + 13, ms.counters.p + 13, ms.from.p,
                            ms.kd.p};
        {
          HL_PROF(P_DEC_MULTI, 0, s);
          hipLaunchKernelGGL(k_acc_degrees, fixed_grid(c), 256, 0, s, u64(c), ms.node.p, ms.state.p, m.g.row_len,
                             ms.mdeg.p);
          HL_LAUNCH_CHECK();
          exclusive_scan_u64_async(ms.mdeg.p, ms.mpre.p, c, tmp, s);
          fill_u32(ms.counters.p + 13, c, 1, s);
        }
        {
          HL_PROF(P_MARK_MOVED, 0, s);
          launch_marks(th, m, fmm, ms.mark.p, nxt, next_cnt, s);
          HL_LAUNCH_CHECK();
        }
        if (m.ck) {
          HL_PROF(P_CACHE_UPDATE, 0, s);
          hipLaunchKernelGGL(k_ck_flat, kMarksBlocks, 256, 0, s, m, fmm);
          HL_LAUNCH_CHECK();
        }
      }
      d2h_group({{&accepted, acc_cnt, sizeof(u32)}, {&next_host, next_cnt, sizeof(u32)}}, s);
      hipLaunchKernelGGL(k_round_reset, 1, 64, 0, s, cur_cnt, ccnt, bcounts, static_cast<u32*>(nullptr), 0u, status);
      HL_LAUNCH_CHECK();
      stream_sync(s);
    }
    ++rounds;
    if (st) st->moved_rounds += accepted;
    if (hook && hook->fn) hook->fn(hook->ctx, nxt, accepted ? next_host : 0u);
    if (accepted == 0) break;  // throughput.rs:177-180 (every accepted gain is > 0)
    std::swap(cur, nxt);
    ci ^= 1;
    host_cnt = next_host;
  }
  return rounds;
}

void refine_proposals(const Csr& g, const u32* f, u32* sm, const double* dS, const u8* blocked, double gamma,
                      const u32* singles, u64 ns, u64 sspace, MoveScratch& ms, const Thresholds& th, u32 maxdeg,
                      const SubRef& sub, hipStream_t s, bool keep_targets) {
  ms.ensure(g.n, 1, s);
  ms.target.reserve(sspace, s);
  if (ms.target_cap != ms.target.cap) {
    HL_CHECK(hipMemsetAsync(ms.target.p, 0, ms.target.cap, s));
    ms.target_cap = ms.target.cap;
  }
  u32* cnt = ms.counters.p + 3;
  u32* ccnt = ms.counters.p + 1;
  hipLaunchKernelGGL(k_round_reset, 1, 64, 0, s, static_cast<u32*>(nullptr), ccnt, ms.counters.p + 4, cnt, u32(ns),
                     static_cast<u32*>(nullptr));
  HL_LAUNCH_CHECK();
  RefinePolicyT pol{f, sm, dS, blocked, gamma};
  propose_round(pol, g, singles, cnt, ns, maxdeg, ms, th, ms.dstat.p + DS_REF, ms.dstat.p + DS_TRY_REF, s);
  CandOut cand{ms.cnode.p, ms.cto.p, ms.cgain.p, ccnt};
  HL_PROF(P_REFINE, 0, s);
  if (keep_targets) {
    hipLaunchKernelGGL(k_ref_targets, fixed_grid(ns), 256, 0, s, cand, ms.target.p);
    hipLaunchKernelGGL(k_ref_apply, fixed_grid(ns), 256, 0, s, cand, ms.target.p, sm, g.k, sub);
    hipLaunchKernelGGL(k_ref_clear_targets, fixed_grid(ns), 256, 0, s, cand, ms.target.p);
  } else {
    hipLaunchKernelGGL(k_ref_apply, fixed_grid(ns), 256, 0, s, cand, nullptr, sm, g.k, sub);
  }
  HL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_round_reset, 1, 64, 0, s, static_cast<u32*>(nullptr), ccnt, ms.counters.p + 4,
                     static_cast<u32*>(nullptr), 0u, static_cast<u32*>(nullptr));
  HL_LAUNCH_CHECK();
}

}  // namespace hl
