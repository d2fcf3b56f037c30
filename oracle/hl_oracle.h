/*
 * hl_oracle.h -- CPU oracle for the HIT-Leiden incremental hierarchical update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU, the reference
 * crate's Algorithm 6 path (src/core/algorithm/hit_leiden.rs:120-161,934-1793,
 * src/core/algorithm/throughput.rs:79-359, src/core/algorithm/parallel_frontier.rs:133-239,
 * src/core/graph/in_memory.rs:30-138).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker / the
 * reported CPU baseline -- never as the product path.
 *
 * Canonical orders (the reference iterates ahash maps, so its tie-breaks, split
 * numbering and CSR row order are not reproducible; these replace them):
 *   - CSR rows sorted by neighbour id; a self-loop is one entry holding the
 *     record weight w and counts as the two entries (w, w) the reference stores.
 *   - argmax over candidate communities / subcommunities in ascending id with
 *     strict '>', the empty community evaluated last with strict '>'.
 *   - k_v = sequential sum over the sorted row; 2m = pairwise (halving) tree
 *     over k_v zero-padded to a power of two.
 *   - community / subcommunity degree totals = per-id sequential sums in vertex order.
 *   - split components numbered by (subcommunity asc, min vertex asc).
 *   - aggregation compress sums per key in emission order; output sorted by key.
 */
#ifndef HL_ORACLE_H
#define HL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mode bits: OR_CONNECTED (D8, D11, D12) and OR_REF_IDS (no D10 relabel: the
   reference's n + v may name a populated community) only with OR_THR;
   OR_AS_WRITTEN (any mode): the reference's full recomputations -- the CSR
   rebuilt through a hash map per batch and level, O(n) scans per refined
   vertex and per candidate subcommunity -- with the same results (the CPU
   baseline's as-written variant) */
enum { OR_DET = 0, OR_THR = 1, OR_CONNECTED = 2, OR_REF_IDS = 4, OR_AS_WRITTEN = 8 };
enum { OR_MAP_F = 0, OR_MAP_G = 1, OR_MAP_S_CUR = 2, OR_MAP_S_PRE = 3 };
enum { OR_SET_A = 0, OR_SET_B = 1, OR_SET_K = 2, OR_SET_R = 3 };

/* ---- whole-state API (mirrors PartitionState + hit_leiden) ---- */
void*    or_create(int levels, double gamma, int mode, int threads);
void     or_destroy(void* h);
/* load = build G^0 from the edges and run Algorithm 6 with dG = all edges on
   the identity state (hit_leiden.rs:3119-3126).  w may be NULL (unit). */
int      or_load(void* h, uint32_t n, uint64_t m, const uint32_t* u, const uint32_t* v, const double* w);
/* build G^0 only, identity maps, no clustering pass */
int      or_load_graph(void* h, uint32_t n, uint64_t m, const uint32_t* u, const uint32_t* v, const double* w);
/* one incremental batch; returns total movement iterations or <0 on error */
int64_t  or_apply(void* h, uint64_t k, const uint32_t* u, const uint32_t* v, const double* alpha);
uint32_t or_level_n(void* h, int level);
int      or_levels(void* h);
int      or_read_map(void* h, int level, int which, uint32_t* out);
int      or_write_map(void* h, int level, int which, const uint32_t* in);
int      or_read_set(void* h, int level, int which, uint8_t* out);
uint32_t or_level_iters(void* h, int level);
/* canonical CSR of level p: offsets[n+1], cols[nnz], w[nnz]; returns nnz (call with NULLs to size) */
uint64_t or_read_graph(void* h, int level, uint64_t* offsets, uint32_t* cols, double* w);
/* the delta applied at level p in the last batch */
uint64_t or_read_delta(void* h, int level, uint32_t* u, uint32_t* v, double* w);
double   or_twice_m(void* h, int level);
/* hierarchy: P rows of n0: rows 0..P-2 = s^p o .. o s^0, row P-1 = f^{P-1} o s^{P-2} o .. o s^0 */
int      or_hierarchy(void* h, uint32_t* out);
/* modularity of a level-0 labelling over G^0 (compute_modularity semantics) */
double   or_modularity_labels(void* h, const uint32_t* labels);

/* ---- function-level API (unit-test entry points on edge lists) ---- */
int or_apply_delta_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                      uint64_t k, const uint32_t* du, const uint32_t* dv, const double* dw, uint32_t n_new,
                      uint64_t* out_m, uint32_t* ou, uint32_t* ov, double* ow);
int or_inc_movement_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                       uint32_t* f, const uint32_t* s, uint64_t k, const uint32_t* du, const uint32_t* dv,
                       const double* dw, double gamma, int mode, uint8_t* outB, uint8_t* outK);
int or_inc_refinement_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                         const uint32_t* f, uint32_t* s, const uint8_t* K, double gamma, int mode, uint8_t* outR);
int64_t or_inc_aggregation_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                              uint64_t k, const uint32_t* du, const uint32_t* dv, const double* dw,
                              uint32_t* s_pre, const uint32_t* s_cur, const uint8_t* R,
                              uint32_t* ou, uint32_t* ov, double* ow, uint64_t cap, uint32_t* next_n);
int or_def_update_fn(int P, const uint32_t* sizes, uint32_t* f_flat, const uint32_t* s_flat, uint8_t* B_flat);
double or_modularity_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                        const uint32_t* labels);
uint64_t or_select_decoupled_fn(uint64_t k, const uint32_t* node, const uint32_t* from, const uint32_t* to,
                                const double* gain, uint32_t* out_nodes);
void or_canonicalize_fn(uint32_t n, uint32_t* labels);
/* test hook: throughput moves onto an empty-community id whose total was not
   zero (counted since the last reset) */
uint64_t or_empty_collisions(int reset);
/* run() / multilevel_leiden (hit_leiden.rs:30-118, 167-323) with the queue
   shuffle seeded (StdRng::seed_from_u64(seed)): first-seen labels, up to
   hier_rows hierarchy rows of n (first-seen per row), *rows = rows produced;
   returns the movement call count, -1 for bad input */
int64_t  or_static_run(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew, double gamma,
                       uint32_t max_levels, uint64_t seed, uint32_t* labels, uint32_t* hier, uint32_t hier_rows,
                       uint32_t* rows);
/* single_level_movement (hit_leiden.rs:764-932, the queue shuffled by a fresh
   StdRng(seed)) and refine_singleton_merge (:487-710) on their own */
int      or_static_movement_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                               uint32_t* f, double gamma, uint64_t seed);
int      or_static_refine_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                             const uint32_t* f, double gamma, uint32_t* sub);
int  or_should_skip_aggregation_fn(uint64_t delta_len, uint32_t n, const uint8_t* R);

#ifdef __cplusplus
}
#endif
#endif
