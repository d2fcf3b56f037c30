// hl_kernels.h -- launchers of the gfx950 kernels (defined in hl_kernels.hip).
#pragma once
#include "hl_common.h"

namespace hl {

// Level CSR view: rows sorted by neighbour, self-loop stored once with its
// record weight (it counts twice in every row sum).
struct Csr {
  u32 n;
  const u64* row_off;
  const u32* row_len;
  const u32* col;
  const double* w;
  const double* k;        // weighted degrees
  const double* twice_m;  // device scalar, or per graph when gid is set
  const u32* gid = nullptr;  // graph of each vertex (several independent graphs in one engine)
};
// 2m of the graph vertex v belongs to
__device__ __forceinline__ double tm_of(const Csr& g, u32 v) { return g.gid ? g.twice_m[g.gid[v]] : g.twice_m[0]; }
// Row entries (col, w) that a kernel reads once, streamed (non-temporal) so
// that they do not evict the label arrays the same kernel gathers at random.
template <class T>
__device__ __forceinline__ T ld_row(const T* p) {
  return __builtin_nontemporal_load(p);
}

// scratch allocator used by the launchers for library temporaries (rocprim)
struct Temp {
  DevBuf<u8> buf;
  void* get(size_t bytes, hipStream_t s) {
    buf.reserve(bytes, s);
    return buf.p;
  }
};

// ---- generic ----
void fill_u8(u8* p, u8 v, u64 n, hipStream_t s);
void fill_u32(u32* p, u32 v, u64 n, hipStream_t s);
void fill_f64(double* p, double v, u64 n, hipStream_t s);
void iota_u32(u32* p, u32 base, u64 n, hipStream_t s);
void copy_u8(u8* dst, const u8* src, u64 n, hipStream_t s);
u64 count_u8(const u8* a, u64 n, DevBuf<u64>& cnt, hipStream_t s);
// nonzero counts of up to four byte arrays (null entries count 0), one sync
void count4_u8(const u8* a0, const u8* a1, const u8* a2, const u8* a3, u64 n, DevBuf<u64>& cnt, u64 out[4],
               hipStream_t s);
// the same into out[0..4) on the device, without synchronising
void count4_u8_async(const u8* a0, const u8* a1, const u8* a2, const u8* a3, u64 n, u64* out, hipStream_t s,
                     bool cleared = false);
void fill4_u8(u8* a0, u8* a1, u8* a2, u8* a3, u8 v, u64 n, hipStream_t s);
// *out = max(*out, max(a[0..n)))  (out must be initialised)
void max_u32(const u32* a, u64 n, u32* out, hipStream_t s);
// out[i] = max(out[i], max of a[i]) for the non-null a[i], one launch
void max_u32_multi(const u32* const a[4], u64 n, u32* const out[4], hipStream_t s);
// indices i with flags[i] != 0, ascending; returns the count (synchronises)
u64 select_flagged(const u8* flags, u64 n, u32* out, Temp& tmp, DevBuf<u64>& cnt, hipStream_t s);
// the same without the read-back: the count lands in *cnt_dev
void select_flagged_async(const u8* flags, u64 n, u32* out, Temp& tmp, u64* cnt_dev, hipStream_t s);
// exclusive scan of u64 values, writes n+1 entries (last = total); returns total
u64 exclusive_scan_u64(const u64* in, u64* out, u64 n, Temp& tmp, hipStream_t s);
// the same without synchronising: the total is left in out[n]
void exclusive_scan_u64_async(const u64* in, u64* out, u64 n, Temp& tmp, hipStream_t s);
// stable sort of (key, value) pairs by key (keys_out/vals_out separate)
void sort_pairs_u64(const u64* kin, u64* kout, const u32* vin, u32* vout, u64 n, int end_bit, Temp& tmp,
                    hipStream_t s);
// stable radix sort of (key, f64 value) pairs
void sort_pairs_u64_f64(const u64* kin, u64* kout, const double* vin, double* vout, u64 n, int end_bit, Temp& tmp,
                        hipStream_t s);
void sort_pairs_u32(const u32* kin, u32* kout, const u32* vin, u32* vout, u64 n, int end_bit, Temp& tmp,
                    hipStream_t s);

// ---- graph maintenance ----
struct ApplyScratch {
  DevBuf<u64> keys, keys_s, ukey, ukey_s, newcap, newoff, chunks, chunkbase, cnt, ucode, uscan;
  DevBuf<u32> chunk_seg;  // merge block -> its segment
  // rows updated in place: their old entries staged here first (offsets by a
  // scan of the staged lengths)
  DevBuf<u64> stlen, stoff;
  DevBuf<u32> st_col;
  DevBuf<double> st_w;
  DevBuf<u32> idx, idx_s, runs, uidx, perm, segs, seg_row, seg_len, new_len, ins_pre, del_pre, seg_ins, seg_del;
  DevBuf<u8> heads, uop;
  DevBuf<double> uw;
  DevBuf<u32> ucount;
  DevBuf<u64> c1;
  DevBuf<u32> dhubs;  // rows whose degree is summed in pieces, then their count
  DevBuf<double> dacc;
};

struct LevelCsrMut {
  u32 n;
  u64* row_off;
  u32* row_len;
  u32* row_cap;
  u32* col;
  double* w;
  double* k;
  double* twice_m;
};

// Maintained totals a kernel keeps current when it changes k or a label
// (null pointers: not maintained).  tot[c] = sum of k over f = c,
// dS[S] = sum of k over s = S, sizes[S] = |{v : s(v) = S}|.
struct TotalsRef {
  double* tot = nullptr;
  const u32* f = nullptr;
  double* dS = nullptr;
  u32* sizes = nullptr;
  const u32* s = nullptr;
  double* tm = nullptr;       // per-graph 2m, kept current from the row sums
  const u32* gid = nullptr;
};

// Plans and applies G <- G (+) dG (in_memory.rs:103-138).  Returns the number of
// affected rows and writes into *need the pool entries the relocation needs;
// the engine must call apply_delta_commit after making room.
struct ApplyPlan {
  // need: pool entries for the rows that outgrow their capacity (moved, with
  // slack); staged: old entries of the rows updated in place
  u64 nruns = 0, nupd = 0, nseg = 0, need = 0, total_chunks = 0, old_entries = 0, max_row = 0, staged = 0;
  u64 new_entries = 0;  // entries of the updated rows afterwards
};
void apply_delta_plan(const Rec* d, u64 k, const LevelCsrMut& g, ApplyScratch& a, Temp& tmp, ApplyPlan& plan,
                      hipStream_t s);
void apply_delta_commit(const LevelCsrMut& g, u64 pool_base, ApplyScratch& a, const ApplyPlan& plan, bool det,
                        const TotalsRef& tr, hipStream_t s);
// Pool compaction in two steps: compact_plan returns the pool entries the rows
// take (their capacities when caps is set, else their lengths) and leaves the
// new row offsets in off_tmp; compact_copy moves every row into [col2, w2)
// (capacities kept) without waiting for it.
u64 compact_plan(const LevelCsrMut& g, DevBuf<u64>& off_tmp, Temp& tmp, hipStream_t s, bool caps = false);
void compact_copy(const LevelCsrMut& g, const u64* offs, u32* col2, double* w2, hipStream_t s);
// the rows laid out at offs (from compact_plan) into col2/w2, leaving the live rows as they are
void gather_rows(const LevelCsrMut& g, const u64* offs, u32* col2, double* w2, hipStream_t s);
// canonical 2m (pairwise tree); partials needs twice_m_scratch(n) doubles
u64 twice_m_scratch(u32 n);
void twice_m_tree(const double* k, u32 n, double* partials, double* out, hipStream_t s);
void grow_init(u32 n_old, u32 n_new, u32 base_f, u32 base_g, u32 base_s, u32* f, u32* g, u32* s_pre, u32* s_cur,
               u64* row_off, u32* row_len, u32* row_cap, double* k, hipStream_t s);

// ---- movement ----
void seed(const Rec* d, u64 k, const u32* f, const u32* s, u8* A, u8* K, hipStream_t st);
void label_totals_atomic(const u32* lab, const double* k, u32 n, double* tot, hipStream_t s);
// canonical (vertex-order sequential) per-label sums; tot zero-filled to tot_n first
void label_totals_canonical(const u32* lab, const double* k, u32 n, double* tot, u64 tot_n, DevBuf<u32>& ks,
                            DevBuf<u32>& vs, DevBuf<u32>& vidx, DevBuf<u8>& heads, DevBuf<u32>& starts,
                            DevBuf<u64>& cnt, Temp& tmp, hipStream_t s);

struct ProposeScratch {
  DevBuf<u32> lists, counts, bigoff_u32;
  DevBuf<u64> bigsize, bigoff, degsum;
  DevBuf<u32> hkeys;
  DevBuf<double> hvals;
};
struct Thresholds {
  u32 small_deg = 16;
  u32 wave_deg = 512;
  u32 dec_rounds = 48;  // priority rounds of the multi-kernel decoupling before the in-order settle
  u32 marks_block = 256;  // workgroup size of the grid-wide neighbour marks (128, 256 or 512)
  u32 marks_grid = 1024;  // their workgroups
  bool persist = true;    // a level's throughput rounds in one cooperative kernel while they fit it
  u32 persist_blocks = 32;  // its workgroups (a grid barrier costs about 0.13 us per workgroup)
};
// deterministic sweep (hit_leiden.rs:1108-1188) in one workgroup
struct DetMoveScratch {
  DevBuf<u32> bm0, bm1, bm2, gkeys, status;
  DevBuf<double> gvals;
  DevBuf<u64> tot_len, pops;
};
// returns 0 when the sweep finished, 1 when the community-total array needs to
// grow (the kernel stopped early, state is resumable)
int det_move(const Csr& g, u32* f, const u32* s, double* tot, u64 tot_cap, u64* tot_len_dev, const u8* A,
             bool fresh, u8* B, u8* K, double gamma, DetMoveScratch& ds, u32 max_deg, hipStream_t st);

// ---- refinement ----
struct RefineScratch {
  DevBuf<u8> aff, incc, keepflag, single, is_cand, blocked, target;
  DevBuf<u32> vcc, parent, cnt32, roots, roots_s, rank_keys_u32, newid, parentIdx, sizes, rlist, prop, smax;
  DevBuf<u64> best64, rkeys, rkeys_s, cnt;
  DevBuf<double> dS, wS, dpar;
  DevBuf<u32> gk;
  DevBuf<double> gv;
  DevBuf<u32> status;
  DevBuf<u32> mhead, mnext;
  DevBuf<u64> kkeys, kkeys_s;
  DevBuf<u32> rsorted;
  DevBuf<u32> hubs;  // k_ws: hub rows listed for the piece kernel, then their count
  DevBuf<u32> sel, sl, vals;
};
// Subcommunity sizes and degrees a caller keeps current across batches
// (throughput mode): size[S] = |{v : s(v) = S}|, deg[S] = sum of k over s = S.
// Null pointers: recomputed where needed.
struct SubRef {
  u32* size = nullptr;
  double* deg = nullptr;
};

// connected-component split of subcommunities touched by K (hit_leiden.rs:1276-1331).
// s_bound: every label of s is below it.  Returns max(s) + 1 after the split and
// sets *identity when s is the identity map after it (hit_leiden.rs:1333-1339).
// f != nullptr links two vertices only through edges inside their community
// (f(u) = f(v)); nullptr links every same-subcommunity neighbour (hit_leiden.rs:1233-1240)
// first_largest keeps the first largest component (the static path) instead of the last;
// maxdeg bounds the row lengths (rows above 4,096 entries are walked in pieces over the grid)
u64 cc_split(const Csr& g, u32* s, const u8* K, u8* R, u64 s_bound, bool* identity, const SubRef& sub,
             RefineScratch& rs, Temp& tmp, hipStream_t st, const u32* f = nullptr, bool first_largest = false,
             u32 maxdeg = 0);
// size[s(v)] += 1 for the vertices [n0, n1) a level grew by (their k is 0)
void sub_add_vertices(const u32* s, u32 n0, u32 n1, u32* size, hipStream_t st);
// from scratch: size and deg over every vertex (arrays zero-filled first, cap entries)
void sub_recompute(const u32* s, const double* k, u32 n, u32* size, double* deg, u64 cap, hipStream_t st);
// throughput refinement over R; keep_targets: the connected-targets rule (a
// singleton another singleton proposes to join stays put), otherwise every
// snapshot proposal applies (throughput.rs:278-287)
struct MoveScratch;
void refine_parallel(const Csr& g, const u32* f, u32* s, const u8* R, double gamma, u64 comm_space, u64 sn,
                     RefineScratch& rs, ProposeScratch& ps, MoveScratch& mvs, u32 maxdeg, const Thresholds& th,
                     Temp& tmp, const double* dpar_in, const SubRef& sub, hipStream_t st, bool keep_targets = true);
// deterministic refinement (hit_leiden.rs:1359-1413)
void refine_det(const Csr& g, const u32* f, u32* s, const u8* R, double gamma, u64 comm_space, u64 sn,
                RefineScratch& rs, DevBuf<u32>& ks, DevBuf<u32>& vs, DevBuf<u32>& vidx, DevBuf<u8>& heads, DevBuf<u32>& starts,
                DevBuf<u64>& cnt, u32 max_deg, Temp& tmp, hipStream_t st);

// ---- aggregation (hit_leiden.rs:1639-1731) ----
struct AggScratch {
  DevBuf<Rec> H;
  DevBuf<u32> rlist, base32, runs, idx, idx_s, keep_idx;
  DevBuf<u64> cnt64, off64, keys, keys_s, cnt;
  DevBuf<u8> heads, keep;
  DevBuf<u64> c1, c2;
  DevBuf<u32> mx;
  DevBuf<Rec> comp;
};
// returns |dG^{p+1}| written to out (grown as needed); *next_n = max(s_cur u s_pre)+1
u64 aggregate(const Csr& g, const Rec* dG, u64 k, u32* s_pre, const u32* s_cur, const u8* R, DevBuf<Rec>& out,
              u32* next_n, AggScratch& as, Temp& tmp, hipStream_t st);

// ---- def_update (hit_leiden.rs:1733-1793) ----
// f^p(v) = f^{p+1}(s^p(v)) for v in B; tot (may be null) is level p's maintained totals
// f[v] <- the rank of f[v] among the labels in use (all below L), so every
// label ends below their count; tot (tot_cap entries, or null) moves with its
// labels into tot2 (reserved for at least tot_need entries and zeroed here),
// which the caller swaps in.  flags holds the presence bitmap, its word
// popcounts and their prefix.
void rank_labels(u32* f, u32 n, u64 L, const double* tot, u64 tot_cap, u64 tot_need, DevBuf<u64>& flags,
                 DevBuf<double>& tot2, Temp& tmp, hipStream_t s);
// f^p(v) = f^q(s^p(v)) for v in B (every v when B is null), keeping tot current
void def_pull(const u8* B, u32* fp, const u32* fq, const u32* sp, u32 n, double* tot, const double* k,
              hipStream_t s);
void def_push(const u8* Bp, u8* Bq, const u32* sq, u32 nq, u64 np, hipStream_t s);

// ---- the static path: run() / multilevel_leiden (hit_leiden.rs:30-118, 167-323) ----
struct StaticScratch {
  DevBuf<u32> queue, gk, gused, list, order, ckey, starts, first;
  DevBuf<double> gv, cval, wc, ws, vals, vals2, cdeg, dC;
  DevBuf<u8> inq, heads;
  DevBuf<u64> cnt, keys, keys2;
  DevBuf<unsigned long long> moves;
};
// A[v] = 1 for both ends of the records with w > 0 across communities (every
// vertex when k = 0): single_level_movement's active set (:776-787)
void static_active(const Rec* d, u64 k, const u32* f, u32 n, u8* A, hipStream_t s);
// single_level_movement's queue loop (:856-928) in one workgroup from the
// shuffled active list queue_host[0..qlen) (active: its flags); f and the
// community degrees cdeg (indexed by label, labels < n) are updated
void static_movement(const Csr& g, u32* f, double* cdeg, const u32* queue_host, u32 qlen, const u8* active,
                     u32 maxdeg, double gamma, StaticScratch& ss, hipStream_t s);
// refine_singleton_merge's merges (:617-706) over the split-off vertices R, on
// subcommunities sub (size and dS current, sspace their id space), community
// degrees dC indexed by f
void static_refine_merge(const Csr& g, const u32* f, u32* sub, u32* size, double* dS, u64 sspace, const double* dC,
                         const u8* R, u32 maxdeg, double gamma, StaticScratch& ss, Temp& tmp, hipStream_t s);
// aggregate_graph (:751-790): records of d[0..k) summed per (min, max) of the
// remapped ends, in input order per key; out sorted by key; returns its count
u64 static_aggregate(const Rec* d, u64 k, const u32* remap, Rec* out, StaticScratch& ss, Temp& tmp, hipStream_t s);
// out[c] = comm of the first (minimum) vertex v with remap[v] = c
void static_coarse_start(const u32* remap, u32 n, u32 cn, const u32* comm, u32* out, StaticScratch& ss,
                         hipStream_t s);
// comm[v] = cf[remap[v]]
void static_compose(const u32* cf, const u32* remap, u32 n, u32* comm, hipStream_t s);

// ---- modularity (hit_leiden.rs:121-161) ----
// rows above kIntraHub entries are summed by workgroups over pieces of kIntraPiece
// entries, from a list of at most kIntraHubList rows
constexpr u32 kIntraHub = 4096, kIntraPiece = 16384, kIntraHubList = 65536;
// The kernels of compute_modularity over g with labels below L, allocating
// nothing: sig[L] per-label degree sums, acc[2] = (intra weight, sum sig^2/2m),
// hubs[hub_cap + 1] the hub-row list and its count; Q = (acc[0] - acc[1]) / 2m.
void modularity_launch(const Csr& g, const u32* lab, u64 L, u32 maxdeg, double* sig, double* acc, u32* hubs,
                       u32 hub_cap, hipStream_t s);
// maxdeg: an upper bound of the row lengths
double modularity(const Csr& g, const u32* lab, u32 maxdeg, DevBuf<double>& sig, DevBuf<double>& acc, DevBuf<u32>& mx,
                  hipStream_t s);
void gather_u32(u32* cur, const u32* map, u32 n, hipStream_t s);
void canonicalize(u32* lab, u32 n, DevBuf<u32>& first, DevBuf<u8>& flag, DevBuf<u64>& pos, Temp& tmp,
                  hipStream_t s);

// ---- partition comparison (contingency_metrics, hit_leiden_incremental.rs:577-666) ----
// out = {NMI = MI / sqrt(H_a H_b), purity a->b, purity b->a, Jaccard of communities la and lb}
void compare_partitions(const u32* a, const u32* b, u64 n, u32 la, u32 lb, double out[4], Temp& tmp,
                        hipStream_t s);

// ---- several independent graphs in one engine (the batched form) ----
// gid[v] = graph of level-0 vertex v from the G+1 vertex offsets
void gid_from_offsets(const u32* offs, u32 G, u32 n, u32* gid, hipStream_t s);
// cnt[g] = records of graph g (cnt has G + 1 entries)
void graph_counts(const Rec* d, u64 k, const u32* gid, u32 G, u32* cnt, hipStream_t s);
// cnt[g] = vertices of list[0..c) in graph g (cnt has G + 1 entries)
void list_graph_counts(const u32* list, u64 c, const u32* gid, u32 G, u32* cnt, hipStream_t s);
void part_from_counts(const u32* cnt, u32 G, u8* part, hipStream_t s);
// R = all vertices of every taking-part graph whose s is the identity (hit_leiden.rs:1333-1339, per graph)
void identity_graphs_to_R(const u32* s, const u32* gid, u32 n, u32 G, const u8* part, u8* nid, u8* R,
                          hipStream_t st);
// gid_q[s(v)] = gid_p[v] for every v of a known graph (gid_p[v] < G)
void scatter_gid(const u32* s, const u32* gid_p, u32 n, u32 G, u32* gid_q, hipStream_t st);
// compute_modularity (hit_leiden.rs:121-161) of each graph of a level-0 CSR with gid set
void modularity_graphs(const Csr& g, const u32* lab, u32 G, double* out, DevBuf<double>& sig, DevBuf<u32>& cg,
                       DevBuf<double>& acc, DevBuf<u32>& mx, hipStream_t s);

}  // namespace hl
