// hl_move.h -- throughput-mode movement and refinement proposals with
// device-side frontier lists (hl_move.hip).
#pragma once
#include "hl_kernels.h"

namespace hl {

constexpr u32 kDecSmall = 4096;  // candidates decoupled inside one workgroup

// candidate list: (node, target, gain), count on the device
struct CandOut {
  u32* node;
  u32* to;
  double* gain;
  u32* cnt;
};

// accepted moves whose neighbour lists are marked over the grid
struct FlatMarks {
  u32* node;
  u32* to;
  u64* pre;  // exclusive prefix of the list lengths, pre[c] = total
  u32* c;    // device count for k_marks_flat (0: nothing to do)
  u32* ck_c = nullptr;  // device count for the cache updates of k_ck_flat
  u32* from = nullptr;  // source community of each move (hub-table updates)
  // 1 where the moved vertex already moved in an earlier round of this level:
  // its affected (K) marks were set then (they depend on s only, which the
  // rounds do not change), so only its next-active marks are left to do
  u8* kd = nullptr;
};

// Hub tables: for a vertex of degree above the LDS workgroup bucket, the
// weight of its row per neighbour community, built by its first full
// evaluation in the rounds of a level and kept current by the neighbours'
// moves, so that later rounds evaluate it from the table (its distinct
// neighbour communities) instead of its row.  A table is valid while bit v of
// `bits` is set; a new level's rounds clear the bitmap and empty the pool.
// Keys are never removed: a community whose last entry left keeps weight 0
// and is skipped (row weights are > 1e-12, so a live key's sum is never 0 in
// exact arithmetic; with inexact sums a key that lost all its entries keeps a
// residue of a few ulps and acts as a non-neighbour community, whose gain is
// never above the empty community's).  A table that fills up is dropped and
// rebuilt by the next full evaluation.
// one table slot: the key and its weight share a 16-byte slot, so an update
// touches one line per key
struct alignas(16) HtSlot {
  u32 key;
  u32 pad;
  double w;
};
struct HubTables {
  u32* bits = nullptr;   // validity bitmap (n bits); nullptr: tables off
  u64* loc = nullptr;    // first slot of v's table in the pool (bits 0-47) and log2 of its slots (56-63)
  HtSlot* slot = nullptr;  // pool: community id (or kNone) and the weight of v's row entries in it
  unsigned long long* used = nullptr;  // pool slots handed out (in chunks)
  u64 pool = 0;
  __device__ bool valid(u32 v) const { return (bits[v >> 5] >> (v & 31)) & 1u; }
  __device__ void where(u32 v, u64& off, u32& cap) const {
    const u64 x = loc[v];
    off = x & ((u64(1) << 48) - 1);
    cap = 1u << u32(x >> 56);
  }
};

// what a movement round reads and writes at one level
struct MoveCtx {
  Csr g;
  u32* f;
  const u32* s;
  double* tot;  // community totals (sized comm_cap)
  const double* k;
  u8* A;
  u8* B;
  u8* K;
  double gamma;
  u64 comm_cap;
  // several graphs per engine: graph of each vertex, graphs taking part in the
  // batch and their record counts at this level (an empty count activates the
  // whole graph, hit_leiden.rs:1054-1057 per graph)
  const u32* gid = nullptr;
  const u8* part = nullptr;
  const u32* gcnt = nullptr;
  // neighbour-community cache: ck[e] = f[col[e]] for every row entry e of
  // vertex v with col[e] != v, kNone for v's self-loop.  Rebuilt when the
  // rounds of a level start and kept current by the moves, so the proposals
  // read it with their row instead of gathering f at random.  pool: entries
  // the row pool holds (the cache's size); 0 leaves the cache off.
  u32* ck = nullptr;
  u64 pool = 0;
  HubTables ht;
  // fs[v] = (s[v] << 32) | f[v], built when the rounds of a level start and
  // kept current by the moves: the neighbour marks read both in one gather
  u64* fs = nullptr;
};

// throughput movement (parallel_frontier.rs:155-236): self-loops skipped,
// candidates in ascending id, empty community id = n + v (the level's labels
// are rank-compacted below n first, so the id is always free)
struct MovePolicyT {
  const u32* f;
  const double* tot;
  double gamma;
  u32 n;
  const u32* ck = nullptr;  // neighbour-community cache (MoveCtx::ck), or none
  HubTables ht;             // hub tables (MoveCtx::ht), or none
  static constexpr bool kEmpty = true;
  static constexpr bool kTables = true;
  __device__ bool table_ok(u32 v) const { return ht.bits && ht.valid(v); }
  // row entry e of vertex v (neighbour col[e])
  // Stream: the row is read by a whole group of lanes at once (non-temporal
  // loads); a lane walking its own short row reads it cached instead
  template <bool Stream = true>
  __device__ bool accept_e(u32 v, const u32* col, u64 e, u32& key, int& mult) const {
    mult = 1;
    if (ck) {
      key = ck[e];
      return key != kNone;
    }
    const u32 u = Stream ? ld_row(col + e) : col[e];
    if (u == v) return false;
    key = f[u];
    return true;
  }
  __device__ u32 cur_key(u32 v) const { return f[v]; }
  __device__ double cur_total(u32 c) const { return tot[c]; }
  __device__ bool ok(u32) const { return true; }
  __device__ double cand_total(u32 c) const { return tot[c]; }
  __device__ u32 empty_id(u32 v) const { return n + v; }
};

// throughput refinement (try_refinement_move, throughput.rs:291-359):
// same-community neighbours keyed by subcommunity, self-loop entries twice,
// blocked subcommunities skipped, no empty candidate
struct RefinePolicyT {
  const u32* f;
  const u32* s;
  const double* dS;
  const u8* blocked;
  double gamma;
  static constexpr bool kEmpty = false;
  static constexpr bool kTables = false;
  __device__ bool table_ok(u32) const { return false; }
  // Stream: the row is read by a whole group of lanes at once (non-temporal
  // loads); a lane walking its own short row reads it cached instead
  template <bool Stream = true>
  __device__ bool accept_e(u32 v, const u32* col, u64 e, u32& key, int& mult) const {
    const u32 u = Stream ? ld_row(col + e) : col[e];
    if (f[u] != f[v]) return false;
    key = s[u];
    mult = (u == v) ? 2 : 1;
    return true;
  }
  __device__ u32 cur_key(u32 v) const { return s[v]; }
  __device__ double cur_total(u32 c) const { return dS[c]; }
  __device__ bool ok(u32 c) const { return !blocked[c]; }
  __device__ double cand_total(u32 c) const { return dS[c]; }
  __device__ u32 empty_id(u32) const { return kNone; }
};

// Device work counters (monotone u64), read by the host once per batch.
// A proposal block (DS_MOVE for movement, DS_REF for refinement) holds per
// degree bucket b (kBuckets of them): [b] vertices evaluated, [kBuckets + b]
// row entries read, [2 kBuckets + b] distinct candidate communities.
constexpr u32 kBuckets = 6;  // lane, half wavefront, wavefront, LDS workgroup, hub workgroup, giant
enum DevStat : u32 {
  DS_MOVE = 0,
  DS_REF = 3 * kBuckets,
  DS_DEC_CAND = 6 * kBuckets,  // candidates decoupled in one workgroup
  DS_DEC_ACC,                  // of which accepted
  DS_DEC_DEG,                  // row entries scanned for next-active marks
  DS_MULTI_CAND,               // the same for the multi-kernel path
  DS_MULTI_ACC,
  DS_MULTI_DEG,
  // vertices, row entries and candidates that the first-try (small-table)
  // kernel of each bucket evaluated without spilling to its overflow kernel
  DS_TRY_MOVE,
  DS_TRY_REF = DS_TRY_MOVE + 3 * kBuckets,
  // hub-table evaluations: vertices, table slots scanned, live candidates
  DS_HT = DS_TRY_REF + 3 * kBuckets,
  DS_SETTLED = DS_HT + 3,  // candidates the in-order settle of the multi-kernel decoupling decided
  // the persistent rounds (k_move_rounds): vertices, row entries and distinct
  // candidate communities of their proposals, then their decoupling's
  // counters in the order of DS_DEC_CAND .. DS_MULTI_DEG
  DS_PR_PROP,
  DS_PR_DEC = DS_PR_PROP + 3,
  DS_COUNT = DS_PR_DEC + 6
};

struct MoveScratch {
  DevBuf<unsigned long long> dstat;
  DevBuf<u32> mark, list_a, list_b, blists, cnode, cto, counters;
  DevBuf<double> cgain;
  DevBuf<u32> minUto, minUfrom, minAto, minAfrom;
  DevBuf<u32> tk;  // workgroup hub tables
  DevBuf<u32> hovf;  // hubs whose keys overflow the LDS tables
  DevBuf<double> tv;
  DevBuf<u32> k32, k32b, v32, o1, o2, node, from, to, state, acc;
  DevBuf<u64> k64, k64b;
  DevBuf<u32> mnode, mto;
  DevBuf<u64> mpre, mdeg;
  DevBuf<u8> target;
  DevBuf<u32> ck;  // neighbour-community cache of the level in movement
  DevBuf<u32> mfrom;  // source communities of the moves marked over the grid
  DevBuf<u8> mkd, kd;  // repeat-move flags of those moves (FlatMarks::kd), both paths
  DevBuf<u32> und;     // undecided counts of a chunk of decoupling priority rounds
  DevBuf<unsigned long long> pout;  // what the persistent rounds report
  // hub tables (HubTables)
  DevBuf<u32> ht_bits;
  DevBuf<HtSlot> ht_slot;
  DevBuf<u64> ht_loc;
  DevBuf<unsigned long long> ht_used;
  DevBuf<u64> fs;  // (s << 32) | f of the level in movement
  size_t target_cap = 0;
  u32 mark_n = 0;  // vertices the mark bitmap covers
  // pinned copies of two rounds' counters and the events that mark them
  // written: the next round is enqueued before the host waits for this one's
  u32* snap = nullptr;
  hipEvent_t snap_ev[2] = {nullptr, nullptr};
  MoveScratch() = default;
  MoveScratch(const MoveScratch&) = delete;
  MoveScratch& operator=(const MoveScratch&) = delete;
  ~MoveScratch();
  void ensure(u64 n, u64 comm_cap, hipStream_t s);
  void trim();  // release the candidate-sized buffers of the multi-kernel path
  // clear the list-deduplication bitmap `mark` for a new round (a fill kernel)
  void fresh_marks(hipStream_t s);
};

struct MoveStats {
  u64 active = 0, evaluated = 0, moved_rounds = 0;
  u64 decide_rounds = 0;  // priority rounds of the multi-kernel decoupling
  u64 persist_rounds = 0;  // rounds run inside the cooperative kernel
};

// Told the vertices active next after the seeding and after every round (a
// device list and its count; count 0 once the frontier is empty).  The rounds
// wait for it, so it may read the level's maps.
struct RoundHook {
  void (*fn)(void* ctx, const u32* list, u32 count) = nullptr;
  void* ctx = nullptr;
};

// Synchronous rounds (<= max_rounds) of the throughput movement at one level;
// seeds A/K from the delta (hit_leiden.rs:1037-1057).  Returns the round count.
u32 movement_throughput(const MoveCtx& m, const Rec* d, u64 k, u32 maxdeg, int max_rounds, MoveScratch& ms,
                        const Thresholds& th, Temp& tmp, MoveStats* st, hipStream_t s,
                        const RoundHook* hook = nullptr);

// Snapshot proposals of the singleton vertices `singles`, with the
// connected-targets rule when keep_targets; updates s in place.
// sub (may be empty): maintained subcommunity sizes/degrees to update with the moves
void refine_proposals(const Csr& g, const u32* f, u32* s, const double* dS, const u8* blocked, double gamma,
                      const u32* singles, u64 ns, u64 sspace, MoveScratch& ms, const Thresholds& th, u32 maxdeg,
                      const SubRef& sub, hipStream_t st, bool keep_targets = true);

}  // namespace hl
