/*
 * hitleiden.h -- C ABI of the MI355X (gfx950) HIT-Leiden incremental update engine.
 *
 * The reference crate (nameearly/hit-leiden, Rust) exposes this path as Rust
 * functions only; each entry point below names the reference interface it
 * replaces.  Plain pointers and sizes, no C++ or torch types: a Rust
 * build.rs-linked `extern "C"` block binds it as is.
 *
 * Ownership: an hl_graph owns every device allocation of one resident graph and
 * its hierarchy.  Host arrays passed in are borrowed for the duration of the
 * call.  Calls on one handle are synchronous and must not overlap (the
 * reference takes `&mut PartitionState`, hit_leiden.rs:936-941).  Bad ids or
 * non-finite weights return HL_INVALID_INPUT instead of panicking (the
 * reference panics, hit_leiden.rs:1060; `run` returns InvalidInput,
 * hit_leiden.rs:31-43).  Inputs are checked before anything changes, so
 * HL_INVALID_INPUT leaves the handle as it was; after HL_BACKEND or HL_OOM
 * from a call that updates the graph the handle's state is unspecified:
 * destroy it, or rebuild it from a checkpoint (hl_load_graph, hl_restore_level,
 * hl_write_partition).
 */
#ifndef HITLEIDEN_H
#define HITLEIDEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HitLeidenError (src/core/error.rs:19-27) + allocation failure */
typedef enum {
  HL_OK = 0,
  HL_INVALID_INPUT = 1, /* HitLeidenError::InvalidInput */
  HL_BACKEND = 2,       /* HitLeidenError::Backend: a HIP runtime error */
  HL_ACCELERATION = 3,  /* HitLeidenError::Acceleration: no gfx950 device */
  HL_OOM = 4
} hl_status;

/* RunMode (src/core/config.rs:19-23) */
typedef enum {
  HL_MODE_DETERMINISTIC = 0, /* sequential lowest-index sweep (hit_leiden.rs:1108-1188) */
  HL_MODE_THROUGHPUT = 1     /* synchronous rounds + decoupled moves (throughput.rs:80-236),
                                refinement keeps merge targets in place so refined
                                communities stay connected */
} hl_mode;

/* partition maps of one level (PartitionState, src/core/partition/state.rs:19-32) */
typedef enum {
  HL_MAP_F = 0,     /* community_mapping_per_level[p]            f^p  */
  HL_MAP_G = 1,     /* refined_community_mapping_per_level[p]    g^p  */
  HL_MAP_S_CUR = 2, /* current_subcommunity_mapping_per_level[p] s_cur^p */
  HL_MAP_S_PRE = 3  /* previous_subcommunity_mapping_per_level[p] s_pre^p */
} hl_map;

/* vertex sets of the last batch at one level (hit_leiden.rs:951-952,1031-1033) */
typedef enum {
  HL_SET_ACTIVE = 0,   /* A after seeding (hit_leiden.rs:1037-1057) */
  HL_SET_CHANGED = 1,  /* B: moved by inc_movement */
  HL_SET_AFFECTED = 2, /* K: affected for refinement */
  HL_SET_REFINED = 3   /* R: returned by inc_refinement */
} hl_set;

/* which partition hl_modularity scores */
typedef enum {
  HL_PART_COMMUNITY = 0, /* f^{P-1} o s^{P-2} o .. o s^0: top-level communities pulled down */
  HL_PART_REFINED = 1,   /* g^0 = state.node_to_comm (hit_leiden.rs:1016) */
  HL_PART_F0 = 2         /* f^0 as stored after def_update */
} hl_partition;

/* RunConfig (src/core/config.rs:25-51) restricted to this path */
typedef struct {
  double gamma;        /* resolution_parameter (config.rs:36) */
  int32_t mode;        /* hl_mode */
  int32_t levels;      /* P >= 1 (PartitionState.levels, state.rs:42) */
  int32_t max_rounds;  /* throughput movement round cap, 0 = 20 (hit_leiden.rs:1077) */
  int32_t device;      /* HIP device ordinal */
  uint32_t small_degree; /* autotuned bucket thresholds; 0 = built-in defaults */
  uint32_t wave_degree;
  uint32_t load_pass_records; /* hl_load_graphs: edges per loading pass, 0 = 2^26 */
  uint32_t key_cache;  /* throughput movement: cache of each row entry's neighbour community,
                          kept current by the moves; 0 = off, 1 = on */
  uint32_t hub_tables;  /* throughput movement: rows above the wavefront bucket evaluated from
                          per-vertex community tables kept current by the moves; 0 = on, 1 = off
                          (off when key_cache is on) */
  uint32_t decide_rounds; /* throughput movement, more than 4,096 candidates: priority rounds of
                             the decoupling before the undecided rest is settled in priority
                             order by one thread (the same result); 0 = 48, at most 2^20 */
  uint32_t refine_reference; /* throughput mode: 0 = a consistent hierarchy (a singleton another
                                singleton proposes to join stays put, the component split links
                                vertices only through edges inside their community, def_update
                                pulls every vertex below the top level); 1 = the reference's
                                rules (throughput.rs:278-287, hit_leiden.rs:1233-1240,1733-1793) */
  uint32_t marks_block;  /* throughput movement: workgroup size of the grid-wide neighbour marks
                            (rows of the moved vertices), 0 = 256, else 128, 256 or 512 */
  uint32_t marks_grid;   /* their workgroups, 0 = 1,024, at most 65,536 */
  uint32_t persist_rounds; /* throughput movement: a level's rounds in one cooperative kernel (grid barriers
                              between proposals, decoupling and marks, no host round trip) while they fit
                              it -- rows of at most 256 entries, no hub tables, at most 4,096 candidates and
                              16,384 listed vertices a round -- else the per-round launches; 0 = on with 32
                              workgroups, 1 = off, 2..1024 = on with that many workgroups (at most one per
                              CU).  The results are the same either way */
} hl_config;

#define HL_MAX_LEVELS 8

/* per-batch counters (returned by load / apply) */
typedef struct {
  uint64_t iterations;               /* hit_leiden() return value (hit_leiden.rs:1017) */
  uint32_t levels;
  uint32_t level_n[HL_MAX_LEVELS];   /* vertices per level after the batch */
  uint64_t delta_len[HL_MAX_LEVELS]; /* |dG^p| */
  uint64_t active[HL_MAX_LEVELS];    /* |A| after seeding */
  uint64_t evaluated[HL_MAX_LEVELS]; /* vertex evaluations summed over rounds */
  uint64_t moved[HL_MAX_LEVELS];     /* |B| */
  uint64_t affected[HL_MAX_LEVELS];  /* |K| */
  uint64_t refined[HL_MAX_LEVELS];   /* |R| */
  uint64_t rounds[HL_MAX_LEVELS];
  uint64_t bytes_alg;                /* algorithmic HBM bytes of the batch */
  double seconds;                    /* host wall time of the call */
  uint64_t decide_rounds[HL_MAX_LEVELS]; /* throughput: priority rounds of the grid-wide decoupling
                                            (rounds with more than 4,096 candidates) */
  uint64_t settled;                  /* throughput: candidates decided by the in-order settle after
                                        decide_rounds priority rounds */
  uint64_t alloc_count;              /* device allocations the call made (grow-only buffers) */
  uint64_t alloc_bytes;              /* their bytes */
  double alloc_seconds;              /* host wall time in those allocations and the frees they replace */
  uint64_t host_syncs;               /* times the host waited for the device during the call */
  uint64_t persist_rounds;           /* throughput: movement rounds run inside the cooperative kernel
                                        (hl_config.persist_rounds), summed over levels */
} hl_batch_stats;

typedef struct hl_graph hl_graph;

/* PartitionState::identity + RunConfig (state.rs:35-49, config.rs:25-51) */
hl_status hl_create(const hl_config* cfg, hl_graph** out);
void hl_destroy(hl_graph* g);

/* Initial load: G^0 <- edges, then Algorithm 6 with dG = all edges on the
   identity state (the reference tests' load pattern, hit_leiden.rs:3119-3126,
   with the empty level-0 CSR pre-pushed).  w == NULL means unit weights
   (Option<f64>::unwrap_or(1.0), in_memory.rs:49).  stats may be NULL.
   n >= 1: an empty vertex set is HL_INVALID_INPUT (here, in hl_load_graph,
   hl_load_graphs and hl_restore_level). */
hl_status hl_load(hl_graph* g, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, const double* w,
                  hl_batch_stats* stats);

/* InMemoryGraph::from (in_memory.rs:30-73) only: builds G^0, identity maps,
   no clustering pass (for restoring a checkpoint with hl_write_partition). */
hl_status hl_load_graph(hl_graph* g, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        const double* w);

/* hit_leiden(state, dG, gamma, mode) (hit_leiden.rs:936-1018): alpha > 0
   inserts/strengthens, alpha < 0 deletes/weakens; ids must be < n. */
hl_status hl_apply_batch(hl_graph* g, uint64_t k, const uint32_t* src, const uint32_t* dst, const double* alpha,
                         hl_batch_stats* stats);

/* Same as hl_apply_batch with the batch already resident in device memory
   (d_* are device pointers on the handle's device). */
hl_status hl_apply_batch_device(hl_graph* g, uint64_t k, const uint32_t* d_src, const uint32_t* d_dst,
                                const double* d_alpha, hl_batch_stats* stats);

hl_status hl_level_count(const hl_graph* g, int32_t* levels);
hl_status hl_level_size(const hl_graph* g, int32_t level, uint32_t* n);

/* copy-out of state.*_mapping_per_level[level] (state.rs:26-31) */
hl_status hl_read_partition(const hl_graph* g, int32_t level, int32_t which, uint32_t* out, uint64_t cap);
/* copy-in (checkpoint restore / seeded starts): n must equal the level size and
   every label must be below 2^31 (0xFFFFFFFF is the kernels' empty key), else
   HL_INVALID_INPUT with the map unchanged */
hl_status hl_write_partition(hl_graph* g, int32_t level, int32_t which, const uint32_t* in, uint64_t n);

/* Checkpoint restore: rebuild level `level` of a one-graph engine from its
   graph (m records, each undirected edge once, a self-loop once with its record
   weight: the hl_read_level_graph content) and its four maps (n entries each;
   NULL leaves a map at the identity).  Restore levels 0..P-1 in order; level 0
   resets the handle.  No counterpart in the reference, which keeps state in
   memory only (README goal 4 lists persistence).  Labels of the four maps must
   be below 2^31, and a throughput engine holds fewer than 2^31 vertices
   (HL_INVALID_INPUT otherwise). */
hl_status hl_restore_level(hl_graph* g, int32_t level, uint32_t n, uint64_t m, const uint32_t* src,
                           const uint32_t* dst, const double* w, const uint32_t* f, const uint32_t* gmap,
                           const uint32_t* s_pre, const uint32_t* s_cur);

/* last batch's A/B/K/R bitsets at one level, one byte per vertex */
hl_status hl_read_frontier(const hl_graph* g, int32_t level, int32_t which, uint8_t* out, uint64_t cap);

/* G^p as canonical CSR (rows sorted, self-loop = one entry with its record
   weight; InMemoryGraph fields, in_memory.rs:20-28).  Call with NULL arrays to
   get nnz. */
hl_status hl_read_level_graph(const hl_graph* g, int32_t level, uint64_t* offsets, uint32_t* cols, double* w,
                              uint64_t cap, uint64_t* nnz);
/* dG^p of the last batch (the inc_aggregation output consumed by level p) */
hl_status hl_read_level_delta(const hl_graph* g, int32_t level, uint32_t* u, uint32_t* v, double* w, uint64_t cap,
                              uint64_t* k);
/* 2m of level p (in_memory.rs:94-96 x 2) */
hl_status hl_twice_m(const hl_graph* g, int32_t level, double* twice_m);

/* PartitionResult.hierarchy_levels (types.rs:107-110): P rows of n0 labels;
   row p < P-1 is s^p o .. o s^0, row P-1 the communities f^{P-1} o s^{P-2} o .. o s^0.
   canonical != 0 relabels each row first-seen contiguous
   (canonicalize_community_ids_in_place, hit_leiden.rs:358-370). */
hl_status hl_read_hierarchy(const hl_graph* g, uint32_t* out, uint64_t cap, int32_t canonical);

/* compute_modularity (hit_leiden.rs:121-161) over G^0 */
hl_status hl_modularity(const hl_graph* g, int32_t which, double* q);
hl_status hl_modularity_labels(const hl_graph* g, const uint32_t* labels, uint64_t n, double* q);

/* run(graph, config) (hit_leiden.rs:30-118) with multilevel_leiden
   (:167-323): the static multilevel Leiden of the graph (src, dst, w) on n
   vertices (w NULL: unit weights), at the handle's gamma and device; the
   handle's loaded state is neither read nor changed.  max_levels is
   RunConfig.max_iterations (> 0).  The movement queue is shuffled by
   StdRng::seed_from_u64(seed) (rand 0.8.5), one generator for the whole run,
   where the reference draws an unseeded thread_rng (:856-860).  labels[n]:
   the communities, numbered first-seen; hier: up to hier_rows rows of n labels
   (hierarchy_levels); *rows: the row count (NULL allowed), which may exceed
   hier_rows; *calls: the movement calls (iteration_count, NULL allowed). */
hl_status hl_run_static(hl_graph* g, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        const double* w, uint32_t max_levels, uint64_t seed, uint32_t* labels, uint32_t* hier,
                        uint32_t hier_rows, uint32_t* rows, int64_t* calls);

/* hl_modularity_labels without any device allocation: no counterpart in the
   reference; for callers that cannot allocate (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  The labels, the
   per-label sums and the partial sums live in a fixed device workspace of
   about 24.3 MiB per GPU (a file-scope device array, bound by hl_create and
   held until the process ends).  Limits: a one-graph engine of at most 2^21
   vertices and labels below 2^21 (HL_INVALID_INPUT beyond them, before any
   device work).  Every call on one GPU uses the same workspace: calls that
   overlap on one GPU (several threads or handles) race. */
hl_status hl_modularity_labels_noalloc(const hl_graph* g, const uint32_t* labels, uint64_t n, double* q);

/* contingency_metrics (hit_leiden_incremental.rs:577-666, the benchmark's quality
   check) on the device for two labelings a, b of n vertices: out[0] NMI =
   MI / sqrt(H(a) H(b)), out[1] purity a->b, out[2] purity b->a, out[3] the
   Jaccard index of community a_largest of a and b_largest of b. */
hl_status hl_compare_partitions(const hl_graph* g, const uint32_t* a, const uint32_t* b, uint64_t n,
                                uint32_t a_largest, uint32_t b_largest, double* out);

/* ---- Batched form: several independent graphs in one engine ----
   No counterpart in the reference, whose PartitionState holds one graph
   (state.rs:19-64); a caller with many graphs would hold one state per graph
   and call hit_leiden once per graph and batch.  Here graph i owns the vertex
   ids [vertex_offsets[i], vertex_offsets[i+1]) of one engine (throughput mode
   only), every edge and delta record joins two vertices of one graph, and one
   hl_apply_batch / hl_apply_batch_device call applies the records of any
   subset of the graphs in one pass: each graph gets exactly the update
   hit_leiden would give it alone (same partition up to label values; 2m,
   empty-delta activation and the identity rule are per graph), graphs without
   records are left untouched.  hl_read_* return the union; hl_modularity does
   not apply (use hl_modularity_graphs). */
hl_status hl_load_graphs(hl_graph* g, uint32_t ngraphs, const uint32_t* vertex_offsets, uint64_t m,
                         const uint32_t* src, const uint32_t* dst, const double* w, hl_batch_stats* stats);
hl_status hl_graph_count(const hl_graph* g, uint32_t* ngraphs);
/* compute_modularity (hit_leiden.rs:121-161) of every graph's communities, q[0..ngraphs) */
hl_status hl_modularity_graphs(const hl_graph* g, double* q, uint64_t cap);

/* ---- Served form ----
   No counterpart in the reference (README goal 3, a resident service, is
   stated there but not built).  hl_serve runs an open-loop request schedule on
   a several-graph engine: request r becomes visible at reqs[r].arrival seconds
   after the call starts; each pass admits, in arrival order, the oldest
   waiting request of every graph (requests of one graph stay in order, one per
   pass) while the pass holds at most max_pass_records delta records (0 admits
   one request per pass, the plain one-batch-per-call path) and its modelled
   device footprint (384 + 16 P bytes per record) fits the device's free
   memory less 1 GiB, applies them together in one hit_leiden pass and reports
   each request's completion time (done[r], seconds from the start) and pass
   index.  A graph's communities (graph-local ids, the top hierarchy row) are
   final once its frontier at the top level is empty, and are handed out at
   that round of the pass (emitted[r], seconds from the start, <= done[r]):
   to emit when emit is not NULL.  Graphs whose top frontier never empties are
   handed out when the pass ends.  With coalesce != 0 a pass takes every waiting request of a
   graph (in arrival order, within the budget) and applies their records as
   one batch of that graph; their results are the graph after that batch.
   Each graph's result is the one it gets alone: the graphs of one pass do not
   interact. */
typedef struct {
  uint32_t graph;       /* graph index */
  uint64_t k;           /* records */
  const uint32_t* src;  /* graph-local vertex ids */
  const uint32_t* dst;
  const double* alpha;  /* NULL: +1.0 each (insertions) */
  double arrival;       /* seconds after the start of hl_serve */
} hl_request;

typedef struct {
  uint64_t passes;
  uint64_t records;
  uint64_t max_pass_requests;
  double seconds;       /* wall time from the start to the last completion */
  double busy_seconds;  /* time spent inside passes */
  uint64_t max_pass_bytes;        /* modelled device footprint of the largest pass */
  uint64_t footprint_cut_passes;  /* passes the free device memory cut short */
} hl_serve_stats;

typedef void (*hl_emit_fn)(void* ctx, uint64_t request, const uint32_t* labels, uint32_t n);

hl_status hl_serve(hl_graph* g, const hl_request* reqs, uint64_t nreq, uint64_t max_pass_records, int32_t coalesce,
                   double* done, uint64_t* pass_of, double* emitted, hl_emit_fn emit, void* ctx,
                   hl_serve_stats* stats);

/* Per-kernel device timing for roofline accounting.  No counterpart in the
   reference (it profiles out of process with perf/samply, profiling.rs:74-152).
   hl_profile_enable(g, 1) resets and starts recording HIP events around each
   kernel family on the handle's stream (on = 2 + slot records that slot only,
   on = 0 stops); hl_profile_read returns, per slot,
   the kernel name, launch count, summed device milliseconds and the summed
   algorithmic HBM bytes of those launches. */
int32_t hl_profile_slots(void);
hl_status hl_profile_enable(hl_graph* g, int32_t on);
hl_status hl_profile_read(const hl_graph* g, int32_t slot, char* name, uint64_t name_cap, uint64_t* launches,
                          double* ms, uint64_t* bytes);

/* HitLeidenError Display (error.rs:21-26) of the last failing call */
const char* hl_last_error(const hl_graph* g);
/* error text of a failed hl_create (no handle yet) */
const char* hl_create_error(void);

#ifdef __cplusplus
}
#endif
#endif
