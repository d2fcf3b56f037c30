// hl_engine.hip -- host orchestration of Algorithm 6 (hit_leiden.rs:936-1018)
// over device-resident levels, and the extern "C" boundary of include/hitleiden.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hitleiden.h"
#include "hl_kernels.h"
#include "hl_move.h"
#include "hl_rng.h"

namespace hl {
namespace {

struct Level {
  u32 n = 0;
  DevBuf<u64> row_off;
  DevBuf<u32> row_len, row_cap;
  DevBuf<double> k;
  DevBuf<u32> col;
  DevBuf<double> w;
  DevBuf<u32> col_spare;  // the other half of the double-buffered row pool (compaction target)
  DevBuf<double> w_spare;
  u64 pool_used = 0;
  DevBuf<double> twice_m;
  DevBuf<u32> f, g, s_pre, s_cur;
  DevBuf<u8> A, B, K, R;  // last batch, before def_update propagation
  DevBuf<Rec> dG;
  u64 dG_len = 0;
  u32 iters = 0;
  // throughput mode: community totals maintained across batches
  DevBuf<double> tot;
  bool tot_valid = false;
  u64 lab_bound = 0;  // every f label < lab_bound
  u64 s_bound = 0;    // every s_cur / s_pre label < s_bound
  u32 deg_bound = 0;  // every row length <= deg_bound
  // several graphs per engine: graph of each vertex id (the graph count for
  // ids no vertex uses); entries below gid_init are initialised
  DevBuf<u32> gid;
  u64 gid_init = 0;
  // throughput mode: subcommunity sizes and degrees over s_cur labels, kept
  // current by every kernel that changes s_cur or k
  DevBuf<u32> ssz;
  DevBuf<double> sdg;
  bool sub_valid = false;

  Csr csr() const { return Csr{n, row_off.p, row_len.p, col.p, w.p, k.p, twice_m.p, gid.p}; }
  LevelCsrMut mut() { return LevelCsrMut{n, row_off.p, row_len.p, row_cap.p, col.p, w.p, k.p, twice_m.p}; }
};

struct KernelTimer {
  // optional per-stage device timing (HIP events on the engine stream)
  bool on = false;
};

}  // namespace

struct Engine {
  hl_config cfg{};
  int P = 1;
  int mode = HL_MODE_THROUGHPUT;
  hipStream_t s = nullptr;
  std::vector<std::unique_ptr<Level>> L;
  u32 n0 = 0;
  bool loaded = false;
  // several independent graphs in one engine (0: one graph)
  u32 G = 0;
  std::vector<u32> goff;  // G + 1 vertex offsets
  DevBuf<u32> goff_d, gcnt, mcg;
  DevBuf<u8> part, gnid;
  std::string err;
  Thresholds th;

  Profiler prof;
  Temp tmp;
  ApplyScratch as;
  ProposeScratch ps;
  DetMoveScratch dms;
  RefineScratch rs;
  AggScratch ag;
  MoveScratch mvs;
  DevBuf<double> tot, partials, sig, acc, tot2;
  DevBuf<u64> clab;  // rank compaction of labels: presence flags, then ranks
  DevBuf<u64> tot_len, cnt, offtmp, lvcnt;
  DevBuf<u32> active, cto, mx, lab, lab2, first;
  DevBuf<double> cgain;
  DevBuf<u8> next, flag8;
  DevBuf<Rec> dcur, dnext;
  DevBuf<u32> ks, vs, vidx, starts;
  DevBuf<u8> heads;
  std::vector<std::unique_ptr<DevBuf<u8>>> Bp, Rp;
  StaticScratch ss;
  // told the top level's frontier after every round (the served form's early emission)
  const RoundHook* top_hook = nullptr;
  // the refined maps below the top level await their pull
  bool g_stale = false;
  void materialize_g() {
    if (!g_stale) return;
    g_stale = false;
    def_update_maps(true);
  }

  ~Engine() {
    L.clear();
    if (s) (void)hipStreamDestroy(s);
  }

  // maxima of up to four u32 arrays of length n with one synchronisation
  void dmax4(const u32* const a[4], u64 n, u32 out[4]) {
    mx.reserve(4, s);
    fill_u32(mx.p, 0, 4, s);
    u32* const outs[4] = {mx.p, mx.p + 1, mx.p + 2, mx.p + 3};
    max_u32_multi(a, n, outs, s);
    hl::d2h_async(out, mx.p, 4 * sizeof(u32), s);
    hl::stream_wait_reads(s);
  }

  u32 dmax(const u32* a, u64 n) {
    mx.reserve(1, s);
    fill_u32(mx.p, 0, 1, s);
    max_u32(a, n, mx.p, s);
    u32 h = 0;
    hl::d2h_async(&h, mx.p, sizeof(u32), s);
    hl::stream_wait_reads(s);
    return h;
  }

  void reserve_level(Level& lv, u64 n) {
    lv.row_off.reserve(n, s, true);
    lv.row_len.reserve(n, s, true);
    lv.row_cap.reserve(n, s, true);
    lv.k.reserve(n, s, true);
    lv.f.reserve(n, s, true);
    lv.g.reserve(n, s, true);
    lv.s_pre.reserve(n, s, true);
    lv.s_cur.reserve(n, s, true);
    lv.A.reserve(n, s);
    lv.B.reserve(n, s);
    lv.K.reserve(n, s);
    lv.R.reserve(n, s);
    lv.twice_m.reserve(1, s);
  }

  // vertex-space growth of level p with fresh ids for new vertices
  void grow_level(int p, u32 n_new) {
    Level& lv = *L[p];
    const u32 n_old = lv.n;
    if (n_new <= n_old) return;
    u64 bf = n_old, bg = n_old, bs = n_old;
    if (n_old) {
      const u32* arrs[4] = {lv.f.p, lv.g.p, lv.s_cur.p, lv.s_pre.p};
      u32 h[4];
      dmax4(arrs, n_old, h);
      bf = std::max<u64>(bf, u64(h[0]) + 1);
      bg = std::max<u64>(bg, u64(h[1]) + 1);
      bs = std::max<u64>(bs, u64(std::max(h[2], h[3])) + 1);
    }
    if (p + 1 < P) bs = std::max<u64>(bs, L[p + 1]->n);
    if (bf + (n_new - n_old) > 0xFFFFFFF0ull || bs + (n_new - n_old) > 0xFFFFFFF0ull)
      throw InputError("id space exhausted");
    reserve_level(lv, n_new);
    grow_init(n_old, n_new, u32(bf), u32(bg), u32(bs), lv.f.p, lv.g.p, lv.s_pre.p, lv.s_cur.p, lv.row_off.p,
              lv.row_len.p, lv.row_cap.p, lv.k.p, s);
    lv.n = n_new;
    lv.lab_bound = std::max<u64>(lv.lab_bound, bf + (n_new - n_old));
    lv.s_bound = std::max<u64>(lv.s_bound, bs + (n_new - n_old));
    if (multi()) ensure_gid(lv, n_new);
    if (lv.sub_valid) {
      ensure_sub(lv);
      sub_add_vertices(lv.s_cur.p, n_old, n_new, lv.ssz.p, s);  // fresh singletons, k = 0
    }
  }

  // maintained subcommunity sizes/degrees: room for every label below s_bound
  // plus one new split id per vertex; computed from scratch when invalid
  void ensure_sub(Level& lv) {
    const u64 need = std::max<u64>(lv.s_bound, lv.n) + lv.n + 1;
    if (lv.sub_valid && lv.ssz.cap >= need) return;
    const u64 old = lv.ssz.cap;
    lv.ssz.reserve(need, s, lv.sub_valid, 1.25);
    lv.sdg.reserve(lv.ssz.cap, s, lv.sub_valid, 1.0);
    if (lv.sub_valid) {
      fill_u32(lv.ssz.p + old, 0, lv.ssz.cap - old, s);
      fill_f64(lv.sdg.p + old, 0.0, lv.sdg.cap - old, s);
    } else {
      sub_recompute(lv.s_cur.p, lv.k.p, lv.n, lv.ssz.p, lv.sdg.p, lv.ssz.cap, s);
      lv.sub_valid = true;
    }
  }
  SubRef sub_of(Level& lv) { return lv.sub_valid ? SubRef{lv.ssz.p, lv.sdg.p} : SubRef{}; }

  bool multi() const { return G > 0; }
  void ensure_gid(Level& lv, u64 need) {
    if (need <= lv.gid_init) return;
    lv.gid.reserve(need, s, true, 1.25);
    fill_u32(lv.gid.p + lv.gid_init, G, lv.gid.cap - lv.gid_init, s);
    lv.gid_init = lv.gid.cap;
  }

  bool par() const { return mode == HL_MODE_THROUGHPUT; }
  // throughput mode with a consistent hierarchy (hl_config.refine_reference == 0)
  bool connected() const { return par() && cfg.refine_reference == 0; }


  // Release scratch sized by the largest call so far (the initial load is a
  // whole-graph batch); it grows back to batch size on demand.
  template <class T>
  static void renew(T& x) {
    x.~T();
    new (&x) T();
  }
  void trim() {
    stream_sync(s);
    for (auto& l : L) {
      l->col_spare.release();
      l->w_spare.release();
    }
    renew(as);
    renew(ag);
    renew(rs);
    renew(ps);
    renew(tmp);
    renew(dnext);
    mvs.trim();
  }

  // make level p's maintained community totals cover every label
  void ensure_totals(Level& lv) {
    const u64 need = std::max<u64>(std::max<u64>(2ull * lv.n, lv.lab_bound), 1) + 1;
    const u64 old = lv.tot.cap;
    if (!lv.tot_valid || old < need) {
      lv.tot.reserve(need, s, lv.tot_valid, 1.25);
      if (lv.tot_valid) {
        fill_f64(lv.tot.p + old, 0.0, lv.tot.cap - old, s);
      } else {
        fill_f64(lv.tot.p, 0.0, lv.tot.cap, s);
        label_totals_atomic(lv.f.p, lv.k.p, lv.n, lv.tot.p, s);
        lv.tot_valid = true;
      }
    }
  }

  // G^p <- G^p (+) dG^p, then 2m
  void apply_delta_level(Level& lv, const Rec* d, u64 k, hl_batch_stats* st, int p) {
    ApplyPlan plan;
    LevelCsrMut g = lv.mut();
    apply_delta_plan(d, k, g, as, tmp, plan, s);
    if (plan.nseg) {
      if (lv.pool_used + plan.need > lv.col.cap) {
        // compact the rows (with their capacities) into the spare pool (kept
        // across batches, so no allocation once it has grown) with room for
        // this batch and headroom: an eighth of the live capacity, or four more
        // batches that move as many rows as this one, within the live capacity
        // (rows that fit their capacity are updated in place and take none)
        const u64 live = compact_plan(g, offtmp, tmp, s, true);
        const u64 room = std::max(live / 8, std::min(4 * plan.need, live));
        const u64 want = std::max<u64>(live + plan.need + room, 1024);
        if (lv.col_spare.cap < live + plan.need + room / 2) {
          stream_sync(s);  // the spare pool may be the source of the last compaction
          lv.col_spare.release();
          lv.w_spare.release();
          lv.col_spare.reserve(want, s, false, 1.0);
          lv.w_spare.reserve(want, s, false, 1.0);
        }
        compact_copy(g, offtmp.p, lv.col_spare.p, lv.w_spare.p, s);
        lv.col.swap(lv.col_spare);
        lv.w.swap(lv.w_spare);
        lv.pool_used = live;
        g = lv.mut();
      }
      TotalsRef tr;
      if (par() && lv.tot_valid) {
        ensure_totals(lv);  // fresh labels of grown vertices
        tr.tot = lv.tot.p;
        tr.f = lv.f.p;
      }
      if (multi()) {
        tr.tm = lv.twice_m.p;  // each graph's 2m follows its rows' degree changes
        tr.gid = lv.gid.p;
      }
      if (par() && lv.sub_valid) {
        tr.dS = lv.sdg.p;
        tr.s = lv.s_cur.p;
      }
      apply_delta_commit(g, lv.pool_used, as, plan, mode == HL_MODE_DETERMINISTIC, tr, s);
      lv.pool_used += plan.need;
      lv.deg_bound = std::max<u32>(lv.deg_bound, u32(plan.max_row));
    }
    if (!multi()) {
      partials.reserve(twice_m_scratch(lv.n), s);
      twice_m_tree(lv.k.p, lv.n, partials.p, lv.twice_m.p, s);
    }
    if (st && p < HL_MAX_LEVELS)
      st->bytes_alg += 16 * k + 24 * plan.old_entries + 12 * plan.new_entries + 20 * plan.nseg;
  }

  u64 comm_space(Level& lv) {
    if (par() && lv.tot_valid) return std::max<u64>(std::max<u64>(2ull * lv.n, lv.lab_bound), 1);
    u64 mf = lv.n ? u64(dmax(lv.f.p, lv.n)) + 1 : 0;
    return std::max<u64>(std::max<u64>(2ull * lv.n, mf), 1);
  }

  // Rank-compact the level's community labels before its throughput rounds
  // (order preserved, so no tie decision changes): afterwards every label is
  // below the number of communities <= n, and the reference's empty ids n + v
  // (parallel_frontier.rs:214) never name a populated community.  Without it,
  // labels that came from elsewhere (fresh ids of grown vertices, labels
  // pulled down from the level above, ids of earlier empty moves at a smaller
  // n) can equal n + v.  The maintained totals move with their labels.
  void compact_labels(Level& lv) {
    const u32 n = lv.n;
    if (n == 0) return;
    const u64 L = std::max<u64>(lv.lab_bound, n);
    if (lv.tot_valid) ensure_totals(lv);
    // the totals' need once the rounds take the empty-community ids n + v
    rank_labels(lv.f.p, n, L, lv.tot_valid ? lv.tot.p : nullptr, lv.tot_valid ? lv.tot.cap : 0, 2ull * n + 1, clab,
                tot2, tmp, s);
    if (lv.tot_valid) lv.tot.swap(tot2);
    lv.lab_bound = n;
  }

  // upper bound of the row lengths (rows only shrink below it)
  u32 max_degree(Level& lv) { return lv.deg_bound; }

  // Algorithm 2 (hit_leiden.rs:1024-1192)
  u32 movement(Level& lv, const Rec* d, u64 k, hl_batch_stats* st, int p) {
    const u32 n = lv.n;
    fill4_u8(lv.A.p, lv.B.p, lv.K.p, lv.R.p, 0, n, s);
    if (k == 0 && !multi()) fill_u8(lv.A.p, 1, n, s);  // empty delta activates every vertex (hit_leiden.rs:1054-1057)
    if (multi()) {
      // records per graph at this level; at level 0 they name the graphs taking part
      graph_counts(d, k, lv.gid.p, G, gcnt.p, s);
      if (p == 0) part_from_counts(gcnt.p, G, part.p, s);
    }
    Csr g = lv.csr();
    if (mode == HL_MODE_DETERMINISTIC) {
      seed(d, k, lv.f.p, lv.s_cur.p, lv.A.p, lv.K.p, s);
      u64 len = std::max<u64>(n, n ? u64(dmax(lv.f.p, n)) + 1 : 0);
      u64 cap = len + n + 1024;
      tot.reserve(cap, s);
      label_totals_canonical(lv.f.p, lv.k.p, n, tot.p, cap, ks, vs, vidx, heads, starts, cnt, tmp, s);
      tot_len.reserve(1, s);
      HL_CHECK(hipMemcpyAsync(tot_len.p, &len, sizeof(u64), hipMemcpyHostToDevice, s));
      const u32 maxd = max_degree(lv);
      bool fresh = true;
      while (true) {
        int r = det_move(g, lv.f.p, lv.s_cur.p, tot.p, cap, tot_len.p, lv.A.p, fresh, lv.B.p, lv.K.p, cfg.gamma,
                         dms, maxd, s);
        if (r == 0) break;
        u64 ncap = cap * 2;
        tot.reserve(ncap, s, true, 1.0);
        fill_f64(tot.p + cap, 0.0, tot.cap - cap, s);
        cap = tot.cap;
        fresh = false;
      }
      if (st && p < HL_MAX_LEVELS) {
        u64 pops = 0;
        hl::d2h_async(&pops, dms.pops.p, sizeof(u64), s);
        hl::stream_wait_reads(s);
        st->evaluated[p] += pops;
        st->rounds[p] = 1;
      }
      return 1;
    }
    // throughput: synchronous rounds, at most max_rounds (hit_leiden.rs:1077-1099)
    compact_labels(lv);
    lv.lab_bound = std::max<u64>(lv.lab_bound, 2ull * n);  // empty-community ids n + v
    ensure_totals(lv);
    const int max_rounds = cfg.max_rounds > 0 ? cfg.max_rounds : 20;
    MoveCtx m{g, lv.f.p, lv.s_cur.p, lv.tot.p, lv.k.p, lv.A.p, lv.B.p, lv.K.p, cfg.gamma, lv.tot.cap};
    // the neighbour-community cache spares the proposals their gathers of f
    // but costs a binary search in the neighbour's row per entry of each
    // moved vertex; on R-MAT scale 24 the searches cost more than the gathers
    // saved, so it is opt-in
    if (cfg.key_cache == 1) m.pool = lv.col.cap;
    // hub tables (on unless turned off): slots for the tables of the rows
    // above the wavefront bucket, at most 2^26 (1 GiB)
    if (cfg.hub_tables == 0 && cfg.key_cache == 0)
      m.ht.pool = std::min<u64>(u64(1) << 26, std::max<u64>(u64(1) << 16, lv.pool_used / 4));
    if (multi()) {
      m.gid = lv.gid.p;
      m.part = part.p;
      m.gcnt = gcnt.p;
    }
    MoveStats ms_st;
    const u32 maxd = max_degree(lv);
    u32 rounds = movement_throughput(m, d, k, maxd, max_rounds, mvs, th, tmp, &ms_st, s,
                                     p == P - 1 ? top_hook : nullptr);
    if (st && p < HL_MAX_LEVELS) {
      st->rounds[p] = rounds;
      st->evaluated[p] += ms_st.evaluated;
      st->decide_rounds[p] += ms_st.decide_rounds;
      st->persist_rounds += ms_st.persist_rounds;
    }
    return rounds;
  }

  // Algorithm 3 (hit_leiden.rs:1265-1416)
  void refinement(Level& lv, hl_batch_stats* st, int p) {
    const u32 n = lv.n;
    if (n == 0) return;
    Csr g = lv.csr();
    bool identity = false;
    if (par()) ensure_sub(lv);
    const u64 sn = cc_split(g, lv.s_cur.p, lv.K.p, lv.R.p, std::max<u64>(lv.s_bound, 1), &identity, sub_of(lv), rs,
                            tmp, s, connected() ? lv.f.p : nullptr, false, max_degree(lv));
    lv.s_bound = std::max(lv.s_bound, sn);
    if (multi())
      identity_graphs_to_R(lv.s_cur.p, lv.gid.p, n, G, part.p, gnid.p, lv.R.p, s);  // per graph
    else if (identity)
      fill_u8(lv.R.p, 1, n, s);
    const u64 cs = comm_space(lv);
    if (mode == HL_MODE_DETERMINISTIC)
      refine_det(g, lv.f.p, lv.s_cur.p, lv.R.p, cfg.gamma, cs, sn, rs, ks, vs, vidx, heads, starts, cnt,
                 max_degree(lv), tmp, s);
    else
      refine_parallel(g, lv.f.p, lv.s_cur.p, lv.R.p, cfg.gamma, cs, sn, rs, ps, mvs, max_degree(lv), th, tmp,
                      lv.tot_valid ? lv.tot.p : nullptr, sub_of(lv), s, connected());
    if (multi() && p + 1 < P) {
      // graphs of the next level's vertices: the graph of any member
      Level& up = *L[p + 1];
      ensure_gid(up, sn);
      // from every vertex, not only the refined ones: a vertex grown at this
      // level keeps its fresh subcommunity id unrefined, and the next level's
      // vertex of that id takes rows (and its graph's 2m) at once
      scatter_gid(lv.s_cur.p, lv.gid.p, n, G, up.gid.p, s);
    }
    (void)st;
  }

  void def_update_maps(bool use_g) {
    // copies of the per-level sets (the reference mutates its bitvecs); a
    // consistent hierarchy pulls every vertex and needs none
    while (int(Bp.size()) < P) Bp.emplace_back(new DevBuf<u8>());
    for (int p = 0; p < P && !connected(); ++p) {
      Level& lv = *L[p];
      Bp[p]->reserve(lv.n + 1, s);
      copy_u8(Bp[p]->p, use_g ? lv.R.p : lv.B.p, lv.n, s);
    }
    for (int p = P - 1; p >= 0; --p) {
      Level& lv = *L[p];
      if (p < P - 1) {
        Level& up = *L[p + 1];
        double* tot = nullptr;
        if (!use_g) {
          lv.lab_bound = std::max(lv.lab_bound, up.lab_bound);  // f^p takes f^{p+1} labels
          if (connected()) {
            // the full pull renames nearly every label (level p's ranks back to
            // the labels of the level above): the totals are rebuilt in one pass
            // when the next rounds start, not kept by two atomics per vertex
            lv.tot_valid = false;
          } else if (par() && lv.tot_valid) {
            ensure_totals(lv);
            tot = lv.tot.p;
          }
        }
        // a consistent hierarchy pulls every vertex (f^p = f^{p+1} o s^p), the
        // reference the changed ones (hit_leiden.rs:1733-1793)
        def_pull(connected() ? nullptr : Bp[p]->p, use_g ? lv.g.p : lv.f.p, use_g ? up.g.p : up.f.p, lv.s_cur.p,
                 lv.n, tot, lv.k.p, s);
      }
      if (p > 0 && !connected()) {
        Level& lo = *L[p - 1];
        def_push(Bp[p]->p, Bp[p - 1]->p, lo.s_cur.p, lo.n, lv.n, s);
      }
    }
  }

  // Algorithmic bytes of the throughput proposal and decoupling kernels from
  // the device work counters (deltas since the previous batch):
  //   movement proposal, per (v, round): 20 + 16 deg(v) + 8 c(v)
  //   refinement proposal, per singleton: 24 + 20 deg(v) + 16 c(v)
  //   decoupling: 20 per candidate, 28 per accepted move, 12 per row entry
  //   scanned for next-active marks
  unsigned long long dlast[DS_COUNT] = {};
  void account_work(const unsigned long long* dnow, hl_batch_stats* st) {
    u64 d[DS_COUNT];
    for (u32 i = 0; i < DS_COUNT; ++i) d[i] = dnow[i] - dlast[i], dlast[i] = dnow[i];
    u64 prop[kBuckets], tried[kBuckets], total = 0;
    for (u32 b = 0; b < kBuckets; ++b) {
      prop[b] = 20 * d[DS_MOVE + b] + 16 * d[DS_MOVE + kBuckets + b] + 8 * d[DS_MOVE + 2 * kBuckets + b] +
                24 * d[DS_REF + b] + 20 * d[DS_REF + kBuckets + b] + 16 * d[DS_REF + 2 * kBuckets + b];
      // the part that the bucket's first-try kernel evaluated
      const u32 tm = DS_TRY_MOVE + 3 * b, tr = DS_TRY_REF + 3 * b;
      tried[b] = 20 * d[tm] + 16 * d[tm + 1] + 8 * d[tm + 2] + 24 * d[tr] + 20 * d[tr + 1] + 16 * d[tr + 2];
      if (tried[b] > prop[b]) tried[b] = prop[b];
      total += prop[b];
    }
    // hub-table evaluations: 20 + 16 per table slot (key and weight) + 8 per live candidate
    const u64 htab = 20 * d[DS_HT] + 16 * d[DS_HT + 1] + 8 * d[DS_HT + 2];
    total += htab;
    const u64 dec = 20 * d[DS_DEC_CAND] + 28 * d[DS_DEC_ACC] + 12 * d[DS_DEC_DEG];
    const u64 multi = 20 * d[DS_MULTI_CAND], apply = 28 * d[DS_MULTI_ACC], marks = 12 * d[DS_MULTI_DEG];
    total += dec + multi + apply + marks;
    // the persistent rounds: their proposals (as a bucket's), decoupling and marks
    const u64 pr = 20 * d[DS_PR_PROP] + 16 * d[DS_PR_PROP + 1] + 8 * d[DS_PR_PROP + 2] + 20 * d[DS_PR_DEC] +
                   28 * d[DS_PR_DEC + 1] + 12 * d[DS_PR_DEC + 2] + 12 * d[DS_PR_DEC + 5];
    total += pr;
    if (st) {
      st->bytes_alg += total;
      st->settled += d[DS_SETTLED];
    }
    if (prof.on) {
      // k_prop2_first also sorts every listed vertex above the lane bucket into
      // its bucket's list: 8 B read (list entry, row length) + 4 B written per
      // such vertex, movement and refinement, hub-table vertices included
      u64 listed = d[DS_HT];
      for (u32 b = 1; b < kBuckets; ++b) listed += d[DS_MOVE + b] + d[DS_REF + b];
      const int slots[15] = {P_PROP_SMALL, P_TRY_HALF, P_TRY_WAVE, P_PROP_WAVE, P_TRY_LDS, P_PROP_LDS, P_TRY_HUB,
                             P_TRY_GIANT, P_PROP_BIG, P_DECOUPLE, P_DEC_MULTI, P_APPLY_MOVES, P_MARK_MOVED,
                             P_PROP_TABLE, P_MOVE_ROUNDS};
      const u64 by[15] = {prop[0] + 12 * listed, tried[1], tried[2], prop[1] + prop[2] - tried[1] - tried[2], tried[3],
                          prop[3] - tried[3], tried[4], tried[5], prop[4] + prop[5] - tried[4] - tried[5],
                          dec, multi, apply, marks, htab, pr};
      for (int i = 0; i < 15; ++i)
        if (prof.records(slots[i])) prof.bytes[slots[i]] += by[i];
    }
  }

  // Algorithm 6 (hit_leiden.rs:936-1018) for a batch already in dcur[0..k)
  u64 run_batch(u64 k, hl_batch_stats* st) {
    lvcnt.reserve(4 * HL_MAX_LEVELS, s);
    // every level's four frontier counts, cleared once for the batch
    if (st) fill_u32(reinterpret_cast<u32*>(lvcnt.p), 0, 8 * HL_MAX_LEVELS, s);
    u64 cur_len = k;
    u32 cur_n = n0;
    u64 total_iters = 0;
    for (int p = 0; p < P; ++p) {
      Level& lv = *L[p];
      grow_level(p, std::max(lv.n, cur_n));
      lv.dG.reserve(cur_len + 1, s);
      if (cur_len)
        HL_CHECK(hipMemcpyAsync(lv.dG.p, dcur.p, cur_len * sizeof(Rec), hipMemcpyDeviceToDevice, s));
      lv.dG_len = cur_len;
      apply_delta_level(lv, dcur.p, cur_len, st, p);
      u32 it = movement(lv, dcur.p, cur_len, st, p);
      lv.iters = it;
      total_iters += it;
      refinement(lv, st, p);
      if (st && p < HL_MAX_LEVELS) {
        // frontier sizes, read back with the batch's last synchronisation
        count4_u8_async(lv.A.p, lv.B.p, lv.K.p, lv.R.p, lv.n, lvcnt.p + 4 * p, s, true);
        st->level_n[p] = lv.n;
        st->delta_len[p] = cur_len;
        st->bytes_alg += 24 * cur_len;
      }
      if (p < P - 1) {
        if (cur_len == 0 && count_u8(lv.R.p, lv.n, cnt, s) == 0) {
          // should_skip_aggregation (hit_leiden.rs:1020-1022)
          u32 a = lv.n ? dmax(lv.s_cur.p, lv.n) : 0;
          u32 b = lv.n ? dmax(lv.s_pre.p, lv.n) : 0;
          cur_n = lv.n ? std::max(a, b) + 1 : 0;
          cur_len = 0;
          continue;
        }
        u32 nn = 0;
        u64 kk = aggregate(lv.csr(), dcur.p, cur_len, lv.s_pre.p, lv.s_cur.p, lv.R.p, dnext, &nn, ag, tmp, s);
        dcur.swap(dnext);
        cur_len = kk;
        cur_n = nn;
        if (st) st->bytes_alg += 80 * kk;
      }
    }
    def_update_maps(false);
    // a consistent hierarchy's refined maps are g^p = g^{p+1} o s^p at every
    // level below the top, which no step of a pass reads: they are pulled when
    // something reads them (materialize_g) instead of after every pass
    if (connected())
      g_stale = true;
    else
      def_update_maps(true);
    unsigned long long dnow[DS_COUNT];
    static_assert(sizeof(dnow) <= hl::kD2HGroupBytes, "work counters exceed one readback group item");
    const bool have_dstat = mvs.dstat.p != nullptr;
    u64 c4[4 * HL_MAX_LEVELS];
    const int np = std::min(P, HL_MAX_LEVELS);
    if (have_dstat && st)
      hl::d2h_group({{dnow, mvs.dstat.p, sizeof(dnow)}, {c4, lvcnt.p, 4 * np * sizeof(u64)}}, s);
    else if (have_dstat)
      hl::d2h_async(dnow, mvs.dstat.p, sizeof(dnow), s);
    else if (st)
      hl::d2h_async(c4, lvcnt.p, 4 * np * sizeof(u64), s);
    hl::stream_sync(s);
    if (have_dstat) account_work(dnow, st);
    if (st)
      for (int p = 0; p < np; ++p) {
        st->active[p] = c4[4 * p];
        st->moved[p] = c4[4 * p + 1];
        st->affected[p] = c4[4 * p + 2];
        st->refined[p] = c4[4 * p + 3];
        st->bytes_alg += 12 * st->moved[p];
      }
    if (st) {
      st->iterations = total_iters;
      st->levels = u32(P);
    }
    return total_iters;
  }

  // ---- the static path: run() / multilevel_leiden (hit_leiden.rs:30-118, 167-323) ----
  // A level holding the graph of records d[0..k) on n vertices (rows as the
  // incremental path builds them, degrees summed in row order, canonical 2m);
  // returns 2m.
  double static_level(Level& lv, u32 n, const Rec* d, u64 k) {
    reserve_level(lv, std::max<u32>(n, 1));
    grow_init(0, n, 0, 0, 0, lv.f.p, lv.g.p, lv.s_pre.p, lv.s_cur.p, lv.row_off.p, lv.row_len.p, lv.row_cap.p,
              lv.k.p, s);
    lv.n = n;
    lv.pool_used = 0;
    lv.deg_bound = 0;
    fill_f64(lv.twice_m.p, 0.0, 1, s);
    if (k) {
      ApplyPlan plan;
      apply_delta_plan(d, k, lv.mut(), as, tmp, plan, s);
      if (plan.nseg) {
        lv.col.reserve(plan.need + 1, s, false, 1.0);
        lv.w.reserve(plan.need + 1, s, false, 1.0);
        apply_delta_commit(lv.mut(), 0, as, plan, true, TotalsRef{}, s);
        lv.pool_used = plan.need;
        lv.deg_bound = u32(plan.max_row);
      }
    }
    partials.reserve(twice_m_scratch(n), s);
    twice_m_tree(lv.k.p, n, partials.p, lv.twice_m.p, s);
    double t = 0.0;
    d2h_async(&t, lv.twice_m.p, sizeof(double), s);
    stream_sync(s);
    return t;
  }

  // single_level_movement (:764-932) on lv from f: the active vertices of the
  // records d[0..k) (all when k = 0) shuffled by rng, then the queue loop
  void static_move(Level& lv, const Rec* d, u64 k, u32* f, StdRng& rng) {
    const u32 n = lv.n;
    if (n == 0) return;
    static_active(d, k, f, n, lv.A.p, s);
    ss.list.reserve(n, s);
    const u64 c = select_flagged(lv.A.p, n, ss.list.p, tmp, ss.cnt, s);
    std::vector<u32> q(c);
    if (c) d2h_async(q.data(), ss.list.p, c * sizeof(u32), s);
    ss.cdeg.reserve(n, s);
    fill_f64(ss.cdeg.p, 0.0, n, s);
    label_totals_atomic(f, lv.k.p, n, ss.cdeg.p, s);  // labels < n
    stream_sync(s);
    rng.shuffle(q.data(), c);
    static_movement(lv.csr(), f, ss.cdeg.p, q.data(), u32(c), lv.A.p, lv.deg_bound, cfg.gamma, ss, s);
    stream_sync(s);  // the queue's host memory is read before it goes
  }

  // refine_singleton_merge (:487-710) of comm on g0 into sub
  void static_refine(Level& g0, double t, const u32* comm, u32* sub) {
    const u32 n = g0.n;
    if (n == 0) return;
    HL_CHECK(hipMemcpyAsync(sub, comm, n * sizeof(u32), hipMemcpyDeviceToDevice, s));
    if (t == 0.0) return;
    fill_u8(g0.K.p, 1, n, s);
    fill_u8(g0.R.p, 0, n, s);
    bool ident = false;
    const u64 sn =
        cc_split(g0.csr(), sub, g0.K.p, g0.R.p, n, &ident, SubRef{}, rs, tmp, s, nullptr, true, g0.deg_bound);
    rs.sizes.reserve(sn, s);
    rs.dS.reserve(sn, s);
    sub_recompute(sub, g0.k.p, n, rs.sizes.p, rs.dS.p, sn, s);
    ss.dC.reserve(n, s);
    fill_f64(ss.dC.p, 0.0, n, s);
    label_totals_atomic(comm, g0.k.p, n, ss.dC.p, s);  // labels < n
    static_refine_merge(g0.csr(), comm, sub, rs.sizes.p, rs.dS.p, sn, ss.dC.p, g0.R.p, g0.deg_bound, cfg.gamma, ss,
                        tmp, s);
  }

  // first-seen relabelling of a[0..n) into out (device) and its label count
  u32 static_first_seen(const u32* a, u32 n, u32* out) {
    if (n == 0) return 0;
    HL_CHECK(hipMemcpyAsync(out, a, n * sizeof(u32), hipMemcpyDeviceToDevice, s));
    DevBuf<u8> fl;
    DevBuf<u64> pos;
    canonicalize(out, n, first, fl, pos, tmp, s);
    return dmax(out, n) + 1;
  }

  // run(): the labels (first-seen) into labels, the hierarchy rows into H;
  // returns the movement call count
  int64_t run_static(u32 n, const Rec* ed_host, u64 m, u32 max_levels, u64 seed, u32* labels,
                     std::vector<std::vector<u32>>& H) {
    StdRng rng(seed);
    Level g0, cg;
    DevBuf<Rec> ed, crec;
    DevBuf<u32> comm, sub, remap, cf, row;
    ed.reserve(m + 1, s);
    if (m) HL_CHECK(hipMemcpyAsync(ed.p, ed_host, m * sizeof(Rec), hipMemcpyHostToDevice, s));
    const double t = static_level(g0, n, ed.p, m);
    const u64 nn = std::max<u32>(n, 1);
    comm.reserve(nn, s);
    sub.reserve(nn, s);
    remap.reserve(nn, s);
    cf.reserve(nn, s);
    row.reserve(nn, s);
    iota_u32(comm.p, 0, n, s);
    auto push_row = [&](const u32* a) {
      std::vector<u32> h(n);
      const u32 c = static_first_seen(a, n, row.p);
      if (n) d2h_async(h.data(), row.p, n * sizeof(u32), s);
      stream_sync(s);
      H.push_back(std::move(h));
      return c;
    };
    int64_t calls = 1;
    static_move(g0, ed.p, m, comm.p, rng);
    static_refine(g0, t, comm.p, sub.p);
    u64 prev = push_row(comm.p);
    for (u32 level = 1; level < max_levels; ++level) {
      const u32 cn = static_first_seen(sub.p, n, remap.p);
      if (cn <= 1) break;
      crec.reserve(m + 1, s);
      const u64 nr = static_aggregate(ed.p, m, remap.p, crec.p, ss, tmp, s);
      if (nr == 0) break;
      DevBuf<u32> rep;
      rep.reserve(cn, s);
      static_coarse_start(remap.p, n, cn, comm.p, rep.p, ss, s);
      static_first_seen(rep.p, cn, cf.p);
      static_level(cg, cn, crec.p, nr);
      static_move(cg, crec.p, nr, cf.p, rng);
      ++calls;
      static_compose(cf.p, remap.p, n, comm.p, s);
      const u64 now = push_row(comm.p);
      static_refine(g0, t, comm.p, sub.p);
      if (now >= prev) break;
      prev = now;
    }
    // final polish on the input graph (:290-305)
    static_move(g0, ed.p, m, comm.p, rng);
    ++calls;
    std::vector<u32> fin(n);
    static_first_seen(comm.p, n, row.p);
    if (n) d2h_async(fin.data(), row.p, n * sizeof(u32), s);
    stream_sync(s);
    if (H.empty() || H.back() != fin) H.push_back(fin);
    std::copy(fin.begin(), fin.end(), labels);
    return calls;
  }

  void reset(u32 n, u64 m_hint) {
    L.clear();
    g_stale = false;
    for (int p = 0; p < P; ++p) L.emplace_back(new Level());
    n0 = n;
    Level& l0 = *L[0];
    grow_level(0, n);
    u64 want = 2 * m_hint + 1024;
    l0.col.reserve(want, s, false, 1.25);
    l0.w.reserve(want, s, false, 1.25);
    l0.pool_used = 0;
    for (int p = 0; p < P; ++p) {
      L[p]->twice_m.reserve(u64(G) + 1, s);
      fill_f64(L[p]->twice_m.p, 0.0, u64(G) + 1, s);
    }
    if (multi()) {
      gcnt.reserve(u64(G) + 1, s);
      part.reserve(u64(G) + 1, s);
      gnid.reserve(u64(G) + 1, s);
      fill_u8(part.p, 0, u64(G) + 1, s);
      goff_d.reserve(u64(G) + 1, s);
      HL_CHECK(hipMemcpyAsync(goff_d.p, goff.data(), (u64(G) + 1) * sizeof(u32), hipMemcpyHostToDevice, s));
      gid_from_offsets(goff_d.p, G, n, l0.gid.p, s);
      stream_sync(s);  // goff's host memory is read before it can change
    }
    loaded = true;
  }

  u32 graph_of(u32 v) const {
    return u32(std::upper_bound(goff.begin(), goff.end(), v) - goff.begin()) - 1;
  }
  void upload_batch(u64 k, const u32* src, const u32* dst, const double* w, u32 n_lim) {
    std::vector<Rec> h(k);
    for (u64 i = 0; i < k; ++i) {
      if (src[i] >= n_lim || dst[i] >= n_lim) throw InputError("endpoint id out of range");
      if (multi() && graph_of(src[i]) != graph_of(dst[i])) throw InputError("record joins two graphs");
      double x = w ? w[i] : 1.0;
      if (!std::isfinite(x)) throw InputError("non-finite weight");
      h[i] = Rec{src[i], dst[i], x};
    }
    dcur.reserve(k + 1, s);
    if (k) HL_CHECK(hipMemcpyAsync(dcur.p, h.data(), k * sizeof(Rec), hipMemcpyHostToDevice, s));
  }

  const Level& level(int p) const {
    if (p < 0 || p >= P) throw InputError("level out of range");
    return *L[p];
  }
  DevBuf<u32>& map_of(Level& lv, int which) {
    if (which == HL_MAP_G) materialize_g();
    switch (which) {
      case HL_MAP_F: return lv.f;
      case HL_MAP_G: return lv.g;
      case HL_MAP_S_CUR: return lv.s_cur;
      case HL_MAP_S_PRE: return lv.s_pre;
      default: throw InputError("unknown map");
    }
  }
  DevBuf<u8>& set_of(Level& lv, int which) {
    switch (which) {
      case HL_SET_ACTIVE: return lv.A;
      case HL_SET_CHANGED: return lv.B;
      case HL_SET_AFFECTED: return lv.K;
      case HL_SET_REFINED: return lv.R;
      default: throw InputError("unknown set");
    }
  }

  // the communities of level-0 vertices [lo, lo + cnt) into out[0..cnt)
  void compose_range(u32 lo, u32 cnt, u32* out) {
    if (cnt == 0) return;
    iota_u32(out, lo, cnt, s);
    for (int p = 0; p < P - 1; ++p) gather_u32(out, L[p]->s_cur.p, cnt, s);
    gather_u32(out, L[P - 1]->f.p, cnt, s);
  }

  // community labels f^{P-1} o s^{P-2} o .. o s^0 (or rows of the hierarchy) into lab
  void compose(int upto, bool with_f, u32* out) {
    iota_u32(out, 0, n0, s);
    for (int p = 0; p < upto; ++p) gather_u32(out, L[p]->s_cur.p, n0, s);
    if (with_f) gather_u32(out, L[upto]->f.p, n0, s);
  }
};

}  // namespace hl

using hl::Engine;

struct hl_graph {
  Engine e;
};

namespace {
thread_local std::string g_create_err;

struct ProfBind {
  explicit ProfBind(hl::Engine* e) { hl::g_prof = e ? &e->prof : nullptr; }
  ~ProfBind() {
    if (hl::g_prof && hl::g_prof->on && !hl::g_prof->pending.empty()) {
      // flush waits for the events the call recorded
      try {
        hl::g_prof->flush();
      } catch (...) {
      }
    }
    hl::g_prof = nullptr;
  }
};

// a call that failed: its staged read-backs point into frames it unwound
void forget_staged(hl_graph* g) {
  if (g && g->e.s) hl::drop_staged(g->e.s);
}

template <class F>
hl_status guard(const hl_graph* gc, F&& fn, bool profile = true) {
  hl_graph* g = const_cast<hl_graph*>(gc);
  try {
    ProfBind pb(g && profile ? &g->e : nullptr);
    hl::g_alloc = hl::AllocStats{};
    fn();
    if (g) g->e.err.clear();
    return HL_OK;
  } catch (const hl::InputError& x) {
    forget_staged(g);
    if (g) g->e.err = std::string("invalid input: ") + x.what();
    return HL_INVALID_INPUT;
  } catch (const hl::HipError& x) {
    forget_staged(g);
    if (g) g->e.err = std::string("backend: ") + x.what();
    return x.code == hipErrorOutOfMemory ? HL_OOM : HL_BACKEND;
  } catch (const std::bad_alloc&) {
    forget_staged(g);
    if (g) g->e.err = "host allocation failed";
    return HL_OOM;
  } catch (const std::exception& x) {
    forget_staged(g);
    if (g) g->e.err = std::string("backend: ") + x.what();
    return HL_BACKEND;
  }
}

constexpr uint64_t kLoadPassRecords = 1ull << 26;  // edges per pass of hl_load_graphs

// ---- the workspace of hl_modularity_labels_noalloc ----
// One file-scope device array per GPU, held until the process ends: the labels
// (copied in), the per-label degree sums, the two partial sums and the hub-row
// list of the modularity kernels.  Host code never uses the array's name as an
// address: hl_create stores hipGetSymbolAddress's pointer per device.
constexpr uint32_t kNoAllocVertices = 1u << 21;  // vertices of the engine's graph
constexpr uint32_t kNoAllocLabels = 1u << 21;    // labels below this
constexpr size_t kScrLab = 0;
constexpr size_t kScrSig = kScrLab + size_t(kNoAllocVertices) * sizeof(uint32_t);
constexpr size_t kScrAcc = kScrSig + size_t(kNoAllocLabels) * sizeof(double);
constexpr size_t kScrHubs = kScrAcc + 256;
constexpr size_t kScratchBytes = kScrHubs + (size_t(hl::kIntraHubList) + 1) * sizeof(uint32_t);
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// the current device's address of `scratch`, taken once per device
hipError_t bind_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_dev[dev] = p;
  return e;
}

// Labels a caller hands in (hl_write_partition, hl_restore_level) stay below
// 2^31: 0xFFFFFFFF is the empty key of every hash table and the no-candidate
// sentinel of the kernels, and a throughput level needs room above its labels
// for the empty-community ids.
constexpr uint32_t kMaxLabel = 0x7FFFFFFFu;
void check_labels(const uint32_t* a, uint64_t n, const char* what) {
  for (uint64_t i = 0; i < n; ++i)
    if (a[i] > kMaxLabel) throw hl::InputError(std::string(what) + " labels must be below 2^31");
}
// throughput empty-community ids n + v need n < 2^31
void check_vertex_count(const hl::Engine& e, uint64_t n) {
  if (e.mode == HL_MODE_THROUGHPUT && n > kMaxLabel)
    throw hl::InputError("throughput mode holds fewer than 2^31 vertices");
}

double now_s() { return hl::host_seconds(); }

// wall time and the call's device allocations into the batch counters
void finish_stats(hl_batch_stats* st, double t0) {
  if (!st) return;
  st->seconds = now_s() - t0;
  st->alloc_count = hl::g_alloc.count;
  st->alloc_bytes = hl::g_alloc.bytes;
  st->alloc_seconds = hl::g_alloc.seconds;
  st->host_syncs = hl::g_alloc.syncs;
}
}  // namespace

extern "C" {

hl_status hl_create(const hl_config* cfg, hl_graph** out) {
  if (!cfg || !out) {
    g_create_err = "null argument";
    return HL_INVALID_INPUT;
  }
  *out = nullptr;
  if (cfg->levels < 1 || cfg->levels > HL_MAX_LEVELS || !(cfg->gamma > 0.0) || !std::isfinite(cfg->gamma) ||
      (cfg->mode != HL_MODE_DETERMINISTIC && cfg->mode != HL_MODE_THROUGHPUT) || cfg->key_cache > 1 ||
      cfg->hub_tables > 1 || cfg->decide_rounds > (1u << 20) || cfg->refine_reference > 1 ||
      (cfg->marks_block != 0 && cfg->marks_block != 128 && cfg->marks_block != 256 && cfg->marks_block != 512) ||
      cfg->marks_grid > 65536 || cfg->persist_rounds > 1024) {
    g_create_err = "invalid config";
    return HL_INVALID_INPUT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    g_create_err = "no HIP device";
    return HL_ACCELERATION;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess ||
      std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_create_err = std::string("device is not gfx950: ") + prop.gcnArchName;
    return HL_ACCELERATION;
  }
  hl_graph* g = new (std::nothrow) hl_graph();
  if (!g) return HL_OOM;
  hl_status st = guard(g, [&] {
    HL_CHECK(hipSetDevice(cfg->device));
    HL_CHECK(bind_scratch(cfg->device));
    HL_CHECK(hipStreamCreateWithFlags(&g->e.s, hipStreamNonBlocking));
    g->e.cfg = *cfg;
    g->e.P = cfg->levels;
    g->e.mode = cfg->mode;
    if (cfg->small_degree) g->e.th.small_deg = cfg->small_degree;
    if (cfg->wave_degree) g->e.th.wave_deg = cfg->wave_degree;
    if (cfg->decide_rounds) g->e.th.dec_rounds = cfg->decide_rounds;
    if (cfg->marks_block) g->e.th.marks_block = cfg->marks_block;
    if (cfg->marks_grid) g->e.th.marks_grid = cfg->marks_grid;
    g->e.th.persist = cfg->persist_rounds != 1;
    if (cfg->persist_rounds > 1) g->e.th.persist_blocks = cfg->persist_rounds;
  });
  if (st != HL_OK) {
    g_create_err = g->e.err;
    delete g;
    return st;
  }
  *out = g;
  return HL_OK;
}

void hl_destroy(hl_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->e.cfg.device);
  delete g;
}

const char* hl_create_error(void) { return g_create_err.c_str(); }

const char* hl_last_error(const hl_graph* g) { return g ? g->e.err.c_str() : "null handle"; }

hl_status hl_load_graph(hl_graph* g, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        const double* w) {
  if (!g || (m && (!src || !dst))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    if (n == 0) throw hl::InputError("a graph needs at least one vertex");
    check_vertex_count(e, n);
    HL_CHECK(hipSetDevice(e.cfg.device));
    e.upload_batch(m, src, dst, w, n);
    e.reset(n, m);
    e.apply_delta_level(*e.L[0], e.dcur.p, m, nullptr, 0);
    hl::stream_sync(e.s);
  });
}

hl_status hl_load(hl_graph* g, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, const double* w,
                  hl_batch_stats* stats) {
  if (!g || (m && (!src || !dst))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    if (n == 0) throw hl::InputError("a graph needs at least one vertex");
    check_vertex_count(e, n);
    HL_CHECK(hipSetDevice(e.cfg.device));
    double t0 = now_s();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    e.G = 0;
    e.goff.clear();
    e.upload_batch(m, src, dst, w, n);
    e.reset(n, m);
    e.run_batch(m, stats);
    e.trim();
    finish_stats(stats, t0);
  });
}

hl_status hl_load_graphs(hl_graph* g, uint32_t ngraphs, const uint32_t* vertex_offsets, uint64_t m,
                         const uint32_t* src, const uint32_t* dst, const double* w, hl_batch_stats* stats) {
  if (!g || !vertex_offsets || ngraphs == 0 || (m && (!src || !dst))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    if (e.mode != HL_MODE_THROUGHPUT) throw hl::InputError("several graphs per engine need the throughput mode");
    if (vertex_offsets[0] != 0) throw hl::InputError("vertex_offsets[0] must be 0");
    for (uint32_t i = 0; i < ngraphs; ++i)
      if (vertex_offsets[i + 1] < vertex_offsets[i]) throw hl::InputError("vertex_offsets must not decrease");
    if (vertex_offsets[ngraphs] == 0) throw hl::InputError("the graphs need at least one vertex");
    check_vertex_count(e, vertex_offsets[ngraphs]);
    HL_CHECK(hipSetDevice(e.cfg.device));
    double t0 = now_s();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    e.G = ngraphs;
    e.goff.assign(vertex_offsets, vertex_offsets + ngraphs + 1);
    const uint32_t n = vertex_offsets[ngraphs];
    // group the edges by graph (counting sort), then load the graphs in passes
    // of at most kLoadPassRecords records: graphs are independent, so each gets
    // the same load as in one pass over everything, with bounded scratch
    std::vector<uint64_t> start(size_t(ngraphs) + 1, 0);
    std::vector<uint32_t> gof(m);
    for (uint64_t i = 0; i < m; ++i) {
      if (src[i] >= n || dst[i] >= n) throw hl::InputError("endpoint id out of range");
      const uint32_t a = e.graph_of(src[i]);
      if (a != e.graph_of(dst[i])) throw hl::InputError("edge joins two graphs");
      if (w && !std::isfinite(w[i])) throw hl::InputError("non-finite weight");
      gof[i] = a;
      ++start[a + 1];
    }
    for (uint32_t i = 0; i < ngraphs; ++i) start[i + 1] += start[i];
    std::vector<hl::Rec> h(m);
    {
      std::vector<uint64_t> pos(start.begin(), start.end() - 1);
      for (uint64_t i = 0; i < m; ++i) h[pos[gof[i]]++] = hl::Rec{src[i], dst[i], w ? w[i] : 1.0};
    }
    std::vector<uint32_t>().swap(gof);
    e.reset(n, m);
    uint32_t g0 = 0;
    do {
      uint32_t g1 = g0 + 1;
      const uint64_t budget = e.cfg.load_pass_records ? e.cfg.load_pass_records : kLoadPassRecords;
      while (g1 < ngraphs && start[g1 + 1] - start[g0] <= budget) ++g1;
      const uint64_t k = start[g1] - start[g0];
      e.dcur.reserve(k + 1, e.s);
      if (k) HL_CHECK(hipMemcpyAsync(e.dcur.p, h.data() + start[g0], k * sizeof(hl::Rec), hipMemcpyHostToDevice, e.s));
      if (stats) std::memset(stats, 0, sizeof(*stats));
      e.run_batch(k, stats);
      g0 = g1;
    } while (g0 < ngraphs);
    e.trim();
    finish_stats(stats, t0);
  });
}

hl_status hl_graph_count(const hl_graph* g, uint32_t* ngraphs) {
  if (!g || !ngraphs) return HL_INVALID_INPUT;
  *ngraphs = g->e.G;
  return HL_OK;
}

namespace {
__global__ void k_pack_batch(const uint32_t* a, const uint32_t* b, const double* w, uint64_t k, uint32_t n,
                             const uint32_t* gid, hl::Rec* out, uint32_t* bad) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < k; i += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t u = a[i], v = b[i];
    double x = w ? w[i] : 1.0;
    if (u >= n || v >= n || !isfinite(x) || (gid && gid[u] != gid[v])) *bad = 1;
    out[i] = hl::Rec{u, v, x};
  }
}
}  // namespace

hl_status hl_apply_batch(hl_graph* g, uint64_t k, const uint32_t* src, const uint32_t* dst, const double* alpha,
                         hl_batch_stats* stats) {
  if (!g || (k && (!src || !dst))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    if (!e.loaded) throw hl::InputError("no graph loaded");
    HL_CHECK(hipSetDevice(e.cfg.device));
    double t0 = now_s();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    e.upload_batch(k, src, dst, alpha, e.n0);
    e.run_batch(k, stats);
    finish_stats(stats, t0);
  });
}

hl_status hl_apply_batch_device(hl_graph* g, uint64_t k, const uint32_t* d_src, const uint32_t* d_dst,
                                const double* d_alpha, hl_batch_stats* stats) {
  if (!g || (k && (!d_src || !d_dst))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    if (!e.loaded) throw hl::InputError("no graph loaded");
    HL_CHECK(hipSetDevice(e.cfg.device));
    double t0 = now_s();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    e.dcur.reserve(k + 1, e.s);
    e.mx.reserve(1, e.s);
    hl::fill_u32(e.mx.p, 0, 1, e.s);
    if (k)
      hipLaunchKernelGGL(k_pack_batch, hl::blocks_for(k), 256, 0, e.s, d_src, d_dst, d_alpha, k, e.n0,
                         e.multi() ? e.L[0]->gid.p : nullptr, e.dcur.p, e.mx.p);
    HL_LAUNCH_CHECK();
    uint32_t bad = 0;
    hl::d2h_async(&bad, e.mx.p, sizeof(uint32_t), e.s);
    hl::stream_wait_reads(e.s);
    if (bad) throw hl::InputError("endpoint id out of range, non-finite weight or a record joining two graphs");
    e.run_batch(k, stats);
    finish_stats(stats, t0);
  });
}

hl_status hl_level_count(const hl_graph* g, int32_t* levels) {
  if (!g || !levels) return HL_INVALID_INPUT;
  *levels = g->e.P;
  return HL_OK;
}

hl_status hl_level_size(const hl_graph* g, int32_t level, uint32_t* n) {
  if (!g || !n) return HL_INVALID_INPUT;
  return guard(g, [&] { *n = g->e.level(level).n; });
}

hl_status hl_read_partition(const hl_graph* g, int32_t level, int32_t which, uint32_t* out, uint64_t cap) {
  if (!g || !out) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    const hl::Level& lc = e.level(level);
    hl::Level& lv = const_cast<hl::Level&>(lc);
    if (cap < lv.n) throw hl::InputError("output buffer too small");
    if (lv.n)
      HL_CHECK(hipMemcpyAsync(out, e.map_of(lv, which).p, lv.n * sizeof(uint32_t), hipMemcpyDeviceToHost, e.s));
    hl::stream_sync(e.s);
  });
}

hl_status hl_write_partition(hl_graph* g, int32_t level, int32_t which, const uint32_t* in, uint64_t n) {
  if (!g || (n && !in)) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    HL_CHECK(hipSetDevice(e.cfg.device));
    hl::Level& lv = const_cast<hl::Level&>(e.level(level));
    if (n != lv.n) throw hl::InputError("length differs from the level size");
    check_labels(in, n, "map");
    if (n) HL_CHECK(hipMemcpyAsync(e.map_of(lv, which).p, in, n * sizeof(uint32_t), hipMemcpyHostToDevice, e.s));
    if (which == HL_MAP_F) {
      // caller-written labels: rebuild the maintained totals on next use
      lv.tot_valid = false;
      lv.lab_bound = n ? std::max<uint64_t>(lv.n, uint64_t(e.dmax(lv.f.p, n)) + 1) : 0;
    } else if ((which == HL_MAP_S_CUR || which == HL_MAP_S_PRE) && n) {
      lv.s_bound = std::max<uint64_t>(lv.s_bound, uint64_t(e.dmax(e.map_of(lv, which).p, n)) + 1);
      if (which == HL_MAP_S_CUR) lv.sub_valid = false;
    }
    hl::stream_sync(e.s);
  });
}

hl_status hl_restore_level(hl_graph* g, int32_t level, uint32_t n, uint64_t m, const uint32_t* src,
                           const uint32_t* dst, const double* w, const uint32_t* f, const uint32_t* gmap,
                           const uint32_t* s_pre, const uint32_t* s_cur) {
  if (!g || (m && (!src || !dst))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    HL_CHECK(hipSetDevice(e.cfg.device));
    if (level < 0 || level >= e.P) throw hl::InputError("level out of range");
    if (level == 0 && n == 0) throw hl::InputError("a graph needs at least one vertex");
    check_vertex_count(e, n);
    {
      const uint32_t* maps[4] = {f, gmap, s_pre, s_cur};
      const char* names[4] = {"f", "g", "s_pre", "s_cur"};
      for (int i = 0; i < 4; ++i)
        if (maps[i]) check_labels(maps[i], n, names[i]);
    }
    if (level == 0) {
      e.G = 0;
      e.goff.clear();
      e.upload_batch(m, src, dst, w, n);
      e.reset(n, m);
    } else {
      if (!e.loaded) throw hl::InputError("restore level 0 first");
      if (e.L[level]->n != 0) throw hl::InputError("level already restored");
      e.upload_batch(m, src, dst, w, n);
      e.grow_level(level, n);
    }
    hl::Level& lv = *e.L[level];
    if (lv.n != n) throw hl::InputError("level size differs");
    e.apply_delta_level(lv, e.dcur.p, m, nullptr, level);
    const uint32_t* maps[4] = {f, gmap, s_pre, s_cur};
    const int which[4] = {HL_MAP_F, HL_MAP_G, HL_MAP_S_PRE, HL_MAP_S_CUR};
    for (int i = 0; i < 4; ++i)
      if (maps[i] && n)
        HL_CHECK(hipMemcpyAsync(e.map_of(lv, which[i]).p, maps[i], n * sizeof(uint32_t), hipMemcpyHostToDevice, e.s));
    hl::stream_sync(e.s);
    // bounds the batches rely on, from the restored labels
    const uint32_t* arrs[4] = {lv.f.p, lv.g.p, lv.s_cur.p, lv.s_pre.p};
    uint32_t mx[4] = {0, 0, 0, 0};
    if (n) e.dmax4(arrs, n, mx);
    lv.tot_valid = false;
    lv.sub_valid = false;
    lv.lab_bound = n ? std::max<uint64_t>(n, uint64_t(mx[0]) + 1) : 0;
    lv.s_bound = n ? std::max<uint64_t>(n, uint64_t(std::max(mx[2], mx[3])) + 1) : 0;
  });
}

hl_status hl_read_frontier(const hl_graph* g, int32_t level, int32_t which, uint8_t* out, uint64_t cap) {
  if (!g || !out) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    hl::Level& lv = const_cast<hl::Level&>(e.level(level));
    if (cap < lv.n) throw hl::InputError("output buffer too small");
    if (lv.n) HL_CHECK(hipMemcpyAsync(out, e.set_of(lv, which).p, lv.n, hipMemcpyDeviceToHost, e.s));
    hl::stream_sync(e.s);
  });
}

hl_status hl_read_level_graph(const hl_graph* g, int32_t level, uint64_t* offsets, uint32_t* cols, double* w,
                              uint64_t cap, uint64_t* nnz) {
  if (!g || !nnz) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    hl::Level& lv = const_cast<hl::Level&>(e.level(level));
    const uint32_t n = lv.n;
    // the rows laid out contiguously in a temporary (the live rows stay as they are)
    hl::LevelCsrMut m = lv.mut();
    hl::DevBuf<uint64_t> offs;
    const uint64_t total = n ? hl::compact_plan(m, offs, e.tmp, e.s) : 0;
    *nnz = total;
    if (!offsets && !cols && !w) return;
    if (!offsets || !cols || !w) throw hl::InputError("need offsets, cols and w");
    if (cap < total) throw hl::InputError("output buffer too small");
    offsets[n] = total;
    if (n == 0) return;
    hl::DevBuf<uint32_t> c2;
    hl::DevBuf<double> w2;
    c2.reserve(total + 1, e.s, false, 1.0);
    w2.reserve(total + 1, e.s, false, 1.0);
    hl::gather_rows(m, offs.p, c2.p, w2.p, e.s);
    HL_CHECK(hipMemcpyAsync(offsets, offs.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, e.s));
    if (total) {
      HL_CHECK(hipMemcpyAsync(cols, c2.p, total * sizeof(uint32_t), hipMemcpyDeviceToHost, e.s));
      HL_CHECK(hipMemcpyAsync(w, w2.p, total * sizeof(double), hipMemcpyDeviceToHost, e.s));
    }
    hl::stream_sync(e.s);
  });
}

hl_status hl_read_level_delta(const hl_graph* g, int32_t level, uint32_t* u, uint32_t* v, double* w, uint64_t cap,
                              uint64_t* k) {
  if (!g || !k) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    hl::Level& lv = const_cast<hl::Level&>(e.level(level));
    *k = lv.dG_len;
    if (!u && !v && !w) return;
    if (!u || !v || !w) throw hl::InputError("need u, v and w");
    if (cap < lv.dG_len) throw hl::InputError("output buffer too small");
    std::vector<hl::Rec> h(lv.dG_len);
    if (lv.dG_len)
      HL_CHECK(hipMemcpyAsync(h.data(), lv.dG.p, lv.dG_len * sizeof(hl::Rec), hipMemcpyDeviceToHost, e.s));
    hl::stream_sync(e.s);
    for (uint64_t i = 0; i < lv.dG_len; ++i) {
      u[i] = h[i].u;
      v[i] = h[i].v;
      w[i] = h[i].w;
    }
  });
}

hl_status hl_twice_m(const hl_graph* g, int32_t level, double* twice_m) {
  if (!g || !twice_m) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    hl::Level& lv = const_cast<hl::Level&>(e.level(level));
    HL_CHECK(hipMemcpyAsync(twice_m, lv.twice_m.p, sizeof(double), hipMemcpyDeviceToHost, e.s));
    hl::stream_sync(e.s);
  });
}

hl_status hl_read_hierarchy(const hl_graph* g, uint32_t* out, uint64_t cap, int32_t canonical) {
  if (!g || !out) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    if (!e.loaded) throw hl::InputError("no graph loaded");
    const uint64_t n = e.n0;
    if (cap < n * uint64_t(e.P)) throw hl::InputError("output buffer too small");
    e.lab.reserve(n + 1, e.s);
    for (int p = 0; p < e.P; ++p) {
      if (p < e.P - 1)
        e.compose(p + 1, false, e.lab.p);
      else
        e.compose(e.P - 1, true, e.lab.p);
      if (canonical) {
        hl::DevBuf<uint8_t> fl;
        hl::DevBuf<uint64_t> pos;
        hl::canonicalize(e.lab.p, uint32_t(n), e.first, fl, pos, e.tmp, e.s);
      }
      if (n) HL_CHECK(hipMemcpyAsync(out + uint64_t(p) * n, e.lab.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, e.s));
    }
    hl::stream_sync(e.s);
  });
}

hl_status hl_modularity(const hl_graph* g, int32_t which, double* q) {
  if (!g || !q) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    if (!e.loaded) throw hl::InputError("no graph loaded");
    e.lab.reserve(e.n0 + 1, e.s);
    if (which == HL_PART_COMMUNITY)
      e.compose(e.P - 1, true, e.lab.p);
    else if (which == HL_PART_REFINED) {
      e.materialize_g();
      HL_CHECK(hipMemcpyAsync(e.lab.p, e.L[0]->g.p, e.n0 * sizeof(uint32_t), hipMemcpyDeviceToDevice, e.s));
    }
    else if (which == HL_PART_F0)
      HL_CHECK(hipMemcpyAsync(e.lab.p, e.L[0]->f.p, e.n0 * sizeof(uint32_t), hipMemcpyDeviceToDevice, e.s));
    else
      throw hl::InputError("unknown partition");
    if (e.multi()) throw hl::InputError("several graphs per engine: use hl_modularity_graphs");
    *q = hl::modularity(e.L[0]->csr(), e.lab.p, e.L[0]->deg_bound, e.sig, e.acc, e.mx, e.s);
  });
}

hl_status hl_modularity_graphs(const hl_graph* g, double* q, uint64_t cap) {
  if (!g || !q) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    if (!e.loaded) throw hl::InputError("no graph loaded");
    if (!e.multi()) throw hl::InputError("the engine holds one graph: use hl_modularity");
    if (cap < e.G) throw hl::InputError("output buffer too small");
    e.lab.reserve(e.n0 + 1, e.s);
    e.compose(e.P - 1, true, e.lab.p);
    hl::modularity_graphs(e.L[0]->csr(), e.lab.p, e.G, q, e.sig, e.mcg, e.acc, e.mx, e.s);
  });
}

namespace {

// Device bytes a served pass holds per record, the model of its admission by
// HBM footprint: the batch and its next-level copy (2 x 16), the level's delta
// copy (16 per level), the row-update plan (canonical keys and their sort,
// about 29 per record, and up to two updates per record with their sort,
// prefixes and segment plan, about 230) and the aggregation's records of the
// next level (2 x 16 per emitted record, bounded here by the record count),
// rounded up.
constexpr uint64_t kServeRecordBytes = 384;
constexpr uint64_t kServeRecordBytesPerLevel = 16;
constexpr uint64_t kServeHeadroom = 1ull << 30;  // device bytes a pass leaves free

// The graphs of one served pass and their requests.  A graph's communities are
// final once its frontier at the top level is empty (the top level's f and
// the s maps below it change no more in the pass), so they are handed out at
// that round, before the pass ends.
struct ServePass {
  Engine* e = nullptr;
  const hl_request* reqs = nullptr;
  hl_emit_fn emit = nullptr;
  void* ctx = nullptr;
  double* emitted = nullptr;
  double t0 = 0.0;
  std::vector<uint32_t> graphs;            // graphs of the pass
  std::vector<std::vector<uint64_t>> of;   // their requests, in arrival order
  std::vector<char> out;                   // handed out
  uint32_t left = 0;
  std::vector<uint32_t> lab, counts;
  hl::DevBuf<uint32_t> dl, dcnt;

  // t: when the graph's communities were final (seconds from the start)
  void hand_out(size_t i, double t) {
    Engine& en = *e;
    const uint32_t gr = graphs[i], lo = en.goff[gr], ng = en.goff[gr + 1] - lo;
    if (emit && ng) {
      dl.reserve(ng, en.s);
      en.compose_range(lo, ng, dl.p);
      lab.resize(ng);
      HL_CHECK(hipMemcpyAsync(lab.data(), dl.p, ng * sizeof(uint32_t), hipMemcpyDeviceToHost, en.s));
      hl::stream_sync(en.s);
    }
    for (uint64_t r : of[i]) {
      if (emitted) emitted[r] = t;
      if (emit) emit(ctx, r, lab.data(), ng);
    }
    out[i] = 1;
    --left;
  }
  void rest(double t) {
    for (size_t i = 0; i < graphs.size(); ++i)
      if (!out[i]) hand_out(i, t);
  }
  static void round(void* c, const uint32_t* list, uint32_t count) {
    ServePass& sp = *static_cast<ServePass*>(c);
    if (sp.left == 0) return;
    const double t = now_s() - sp.t0;
    if (count == 0) {
      sp.rest(t);
      return;
    }
    Engine& en = *sp.e;
    sp.dcnt.reserve(uint64_t(en.G) + 1, en.s);
    hl::list_graph_counts(list, count, en.L[en.P - 1]->gid.p, en.G, sp.dcnt.p, en.s);
    sp.counts.resize(uint64_t(en.G) + 1);
    HL_CHECK(hipMemcpyAsync(sp.counts.data(), sp.dcnt.p, (uint64_t(en.G) + 1) * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, en.s));
    hl::stream_sync(en.s);
    for (size_t i = 0; i < sp.graphs.size(); ++i)
      if (!sp.out[i] && sp.counts[sp.graphs[i]] == 0) sp.hand_out(i, t);
  }
};

}  // namespace

hl_status hl_serve(hl_graph* g, const hl_request* reqs, uint64_t nreq, uint64_t max_pass_records, int32_t coalesce,
                   double* done, uint64_t* pass_of, double* emitted, hl_emit_fn emit, void* ctx,
                   hl_serve_stats* stats) {
  if (!g || (nreq && (!reqs || !done))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    if (!e.loaded || !e.multi()) throw hl::InputError("hl_serve needs graphs loaded with hl_load_graphs");
    for (uint64_t r = 0; r < nreq; ++r) {
      const hl_request& q = reqs[r];
      if (q.graph >= e.G || (q.k && (!q.src || !q.dst)) || !std::isfinite(q.arrival))
        throw hl::InputError("invalid request");
      const uint32_t ng = e.goff[q.graph + 1] - e.goff[q.graph];
      for (uint64_t i = 0; i < q.k; ++i) {
        if (q.src[i] >= ng || q.dst[i] >= ng) throw hl::InputError("request endpoint outside its graph");
        if (q.alpha && !std::isfinite(q.alpha[i])) throw hl::InputError("non-finite weight");
      }
    }
    HL_CHECK(hipSetDevice(e.cfg.device));
    std::vector<uint64_t> order(nreq);
    for (uint64_t r = 0; r < nreq; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint64_t a, uint64_t b) { return reqs[a].arrival < reqs[b].arrival; });
    std::vector<std::vector<uint64_t>> waiting(e.G);  // per graph, in arrival order
    std::vector<size_t> head(e.G, 0);
    std::vector<uint32_t> active;                     // graphs with waiting requests
    std::vector<char> is_active(e.G, 0);
    std::vector<hl::Rec> h;
    hl_serve_stats st{};
    uint64_t next = 0, completed = 0, pass = 0;
    const uint64_t per_record = kServeRecordBytes + kServeRecordBytesPerLevel * uint64_t(e.P);
    const double t0 = now_s();
    ServePass sp;
    sp.e = &e;
    sp.reqs = reqs;
    sp.emit = emit;
    sp.ctx = ctx;
    sp.emitted = emitted;
    sp.t0 = t0;
    const bool early = (emit || emitted) && e.mode == HL_MODE_THROUGHPUT;
    hl::RoundHook hook{&ServePass::round, &sp};
    while (completed < nreq) {
      double now = now_s() - t0;
      while (next < nreq && reqs[order[next]].arrival <= now) {
        const uint64_t r = order[next++];
        const uint32_t gr = reqs[r].graph;
        waiting[gr].push_back(r);
        if (!is_active[gr]) {
          is_active[gr] = 1;
          active.push_back(gr);
        }
      }
      if (active.empty()) {
        const double wait = reqs[order[next]].arrival - now;
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
        continue;
      }
      // admission by HBM footprint: the records whose pass fits the device's
      // free memory (less a headroom), within the record budget when one is set
      size_t free_b = 0, total_b = 0;
      HL_CHECK(hipMemGetInfo(&free_b, &total_b));
      const uint64_t fit = free_b > kServeHeadroom ? (uint64_t(free_b) - kServeHeadroom) / per_record : 0;
      const uint64_t cap = std::min<uint64_t>(max_pass_records, fit);
      // admit the oldest waiting request of each graph, oldest first, within the cap
      std::vector<std::pair<double, uint32_t>> heads;
      for (uint32_t gr : active) heads.push_back({reqs[waiting[gr][head[gr]]].arrival, gr});
      std::sort(heads.begin(), heads.end());
      std::vector<uint64_t> admitted;  // grouped by graph, each graph's requests in arrival order
      sp.graphs.clear();
      sp.of.clear();
      uint64_t recs = 0;
      bool full = false, by_bytes = false;
      for (const auto& hg : heads) {
        const uint32_t gr = hg.second;
        for (size_t j = head[gr]; j < waiting[gr].size() && !full; ++j) {
          const uint64_t r = waiting[gr][j];
          if (!admitted.empty() && (max_pass_records == 0 || recs + reqs[r].k > cap)) {
            full = true;
            by_bytes = max_pass_records != 0 && recs + reqs[r].k > fit;
            break;
          }
          if (sp.graphs.empty() || sp.graphs.back() != gr) {
            sp.graphs.push_back(gr);
            sp.of.emplace_back();
          }
          sp.of.back().push_back(r);
          admitted.push_back(r);
          recs += reqs[r].k;
          if (!coalesce) break;  // one request of this graph per pass
        }
        if (full) break;
      }
      h.resize(recs);
      uint64_t o = 0;
      for (uint64_t r : admitted) {
        const hl_request& q = reqs[r];
        const uint32_t base = e.goff[q.graph];
        for (uint64_t i = 0; i < q.k; ++i) h[o++] = hl::Rec{base + q.src[i], base + q.dst[i], q.alpha ? q.alpha[i] : 1.0};
      }
      sp.out.assign(sp.graphs.size(), 0);
      sp.left = uint32_t(sp.graphs.size());
      const double tp = now_s();
      e.dcur.reserve(recs + 1, e.s);
      if (recs) HL_CHECK(hipMemcpyAsync(e.dcur.p, h.data(), recs * sizeof(hl::Rec), hipMemcpyHostToDevice, e.s));
      e.top_hook = early ? &hook : nullptr;
      try {
        e.run_batch(recs, nullptr);
      } catch (...) {
        e.top_hook = nullptr;
        throw;
      }
      e.top_hook = nullptr;
      const double tdone = now_s() - t0;
      st.busy_seconds += now_s() - tp;
      if (emit || emitted) sp.rest(tdone);  // graphs whose top frontier never emptied: at the pass end
      for (uint64_t r : admitted) {
        const uint32_t gr = reqs[r].graph;
        done[r] = tdone;
        if (pass_of) pass_of[r] = pass;
        if (++head[gr] == waiting[gr].size()) {
          waiting[gr].clear();
          head[gr] = 0;
          is_active[gr] = 0;
        }
      }
      active.erase(std::remove_if(active.begin(), active.end(), [&](uint32_t gr) { return !is_active[gr]; }),
                   active.end());
      completed += admitted.size();
      st.records += recs;
      st.max_pass_requests = std::max<uint64_t>(st.max_pass_requests, admitted.size());
      st.max_pass_bytes = std::max<uint64_t>(st.max_pass_bytes, recs * per_record);
      if (by_bytes) ++st.footprint_cut_passes;
      ++pass;
    }
    st.passes = pass;
    st.seconds = now_s() - t0;
    if (stats) *stats = st;
  });
}

hl_status hl_compare_partitions(const hl_graph* g, const uint32_t* a, const uint32_t* b, uint64_t n,
                                uint32_t a_largest, uint32_t b_largest, double* out) {
  if (!g || !out || (n && (!a || !b))) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    hl::DevBuf<uint32_t> da, db;
    da.reserve(n + 1, e.s);
    db.reserve(n + 1, e.s);
    if (n) {
      HL_CHECK(hipMemcpyAsync(da.p, a, n * sizeof(uint32_t), hipMemcpyHostToDevice, e.s));
      HL_CHECK(hipMemcpyAsync(db.p, b, n * sizeof(uint32_t), hipMemcpyHostToDevice, e.s));
    }
    hl::compare_partitions(da.p, db.p, n, a_largest, b_largest, out, e.tmp, e.s);
    hl::stream_sync(e.s);
  });
}

int32_t hl_profile_slots(void) { return hl::P_COUNT; }

hl_status hl_profile_enable(hl_graph* g, int32_t on) {
  if (!g) return HL_INVALID_INPUT;
  return guard(g, [&] {
    HL_CHECK(hipStreamSynchronize(g->e.s));
    g->e.prof.flush();
    g->e.prof.reset();
    g->e.prof.on = on != 0;
    g->e.prof.only = on >= 2 ? on - 2 : -1;
  });
}

hl_status hl_profile_read(const hl_graph* g, int32_t slot, char* name, uint64_t name_cap, uint64_t* launches,
                          double* ms, uint64_t* bytes) {
  if (!g || slot < 0 || slot >= hl::P_COUNT) return HL_INVALID_INPUT;
  const hl::Profiler& p = g->e.prof;
  if (name && name_cap) {
    std::strncpy(name, hl::prof_name(slot), name_cap - 1);
    name[name_cap - 1] = 0;
  }
  if (launches) *launches = p.launches[slot];
  if (ms) *ms = p.ms[slot];
  if (bytes) *bytes = p.bytes[slot];
  return HL_OK;
}

hl_status hl_modularity_labels(const hl_graph* g, const uint32_t* labels, uint64_t n, double* q) {
  if (!g || !q || (n && !labels)) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = const_cast<Engine&>(g->e);
    HL_CHECK(hipSetDevice(e.cfg.device));
    if (!e.loaded) throw hl::InputError("no graph loaded");
    if (e.multi()) throw hl::InputError("several graphs per engine: use hl_modularity_graphs");
    if (n != e.n0) throw hl::InputError("labels length differs from n");
    e.lab.reserve(n + 1, e.s);
    if (n) HL_CHECK(hipMemcpyAsync(e.lab.p, labels, n * sizeof(uint32_t), hipMemcpyHostToDevice, e.s));
    *q = hl::modularity(e.L[0]->csr(), e.lab.p, e.L[0]->deg_bound, e.sig, e.acc, e.mx, e.s);
  });
}

hl_status hl_run_static(hl_graph* g, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        const double* w, uint32_t max_levels, uint64_t seed, uint32_t* labels, uint32_t* hier,
                        uint32_t hier_rows, uint32_t* rows, int64_t* calls) {
  if (!g || (m && (!src || !dst)) || (n && !labels) || (hier_rows && !hier)) return HL_INVALID_INPUT;
  return guard(g, [&] {
    Engine& e = g->e;
    HL_CHECK(hipSetDevice(e.cfg.device));
    if (max_levels == 0) throw hl::InputError("max_iterations must be > 0");
    if (n > kMaxLabel) throw hl::InputError("the static path holds fewer than 2^31 vertices");
    std::vector<hl::Rec> ed(m);
    for (uint64_t i = 0; i < m; ++i) {
      if (src[i] >= n || dst[i] >= n) throw hl::InputError("edge endpoint exceeds node_count");
      const double x = w ? w[i] : 1.0;
      if (!std::isfinite(x)) throw hl::InputError("non-finite weight");
      ed[i] = hl::Rec{src[i], dst[i], x};
    }
    std::vector<std::vector<uint32_t>> H;
    const int64_t c = e.run_static(n, ed.data(), m, max_levels, seed, labels, H);
    const uint32_t nr = std::min<uint32_t>(uint32_t(H.size()), hier_rows);
    for (uint32_t r = 0; r < nr; ++r) std::copy(H[r].begin(), H[r].end(), hier + uint64_t(r) * n);
    if (rows) *rows = uint32_t(H.size());
    if (calls) *calls = c;
  });
}

hl_status hl_modularity_labels_noalloc(const hl_graph* g, const uint32_t* labels, uint64_t n, double* q) {
  if (!g || !q || (n && !labels)) return HL_INVALID_INPUT;
  return guard(
      g,
      [&] {
        Engine& e = const_cast<Engine&>(g->e);
        if (!e.loaded) throw hl::InputError("no graph loaded");
        if (e.multi()) throw hl::InputError("several graphs per engine: use hl_modularity_graphs");
        if (n != e.n0) throw hl::InputError("labels length differs from n");
        if (n > kNoAllocVertices) throw hl::InputError("more than 2^21 vertices: use hl_modularity_labels");
        uint32_t hm = 0;
        for (uint64_t i = 0; i < n; ++i) hm = std::max(hm, labels[i]);
        if (n && hm >= kNoAllocLabels) throw hl::InputError("labels must be below 2^21: use hl_modularity_labels");
        HL_CHECK(hipSetDevice(e.cfg.device));
        unsigned char* base = static_cast<unsigned char*>(g_scratch_dev[e.cfg.device]);
        if (!base) throw hl::InputError("the device workspace is not bound");
        uint32_t* dlab = reinterpret_cast<uint32_t*>(base + kScrLab);
        double* sig = reinterpret_cast<double*>(base + kScrSig);
        double* acc = reinterpret_cast<double*>(base + kScrAcc);
        uint32_t* hubs = reinterpret_cast<uint32_t*>(base + kScrHubs);
        const hl::Csr csr = e.L[0]->csr();
        double t = 0.0;
        HL_CHECK(hipMemcpyAsync(&t, csr.twice_m, sizeof(double), hipMemcpyDeviceToHost, e.s));
        HL_CHECK(hipStreamSynchronize(e.s));
        if (t == 0.0 || n == 0) {
          *q = 0.0;
          return;
        }
        HL_CHECK(hipMemcpyAsync(dlab, labels, n * sizeof(uint32_t), hipMemcpyHostToDevice, e.s));
        hl::modularity_launch(csr, dlab, uint64_t(hm) + 1, e.L[0]->deg_bound, sig, acc, hubs,
                              std::min<uint32_t>(uint32_t(n), hl::kIntraHubList), e.s);
        double h[2] = {0.0, 0.0};
        HL_CHECK(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, e.s));
        HL_CHECK(hipStreamSynchronize(e.s));
        *q = (h[0] - h[1]) / t;
      },
      false);
}

}  // extern "C"
