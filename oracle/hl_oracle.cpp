// hl_oracle.cpp -- CPU oracle for the HIT-Leiden incremental hierarchical update.
//
// TEST INFRASTRUCTURE ONLY (see hl_oracle.h).  Every function below cites the
// reference (Rust) function it restates.  Deviations from the reference, each
// needed either for reproducibility or to make levels > 1 well defined (the
// reference hard-wires one level, src/core/partition/state.rs:42):
//   D1 community-degree arrays sized max(n, max label + 1)    (hit_leiden.rs:1060 indexes labels >= n)
//   D2 load pre-creates an empty level-0 CSR                  (hit_leiden.rs:944-946 would double weights)
//   D3 apply_delta keeps existing self-loops                   (in_memory.rs:125 drops them)
//   D4 aggregation emits an R-R edge once (v_j not in R or i < j) instead of
//      (s_cur(v_j) = s_pre(v_j) or i < j), which double-emits when an unchanged
//      v_j in R has a smaller id than a changed neighbour   (hit_leiden.rs:1664-1666)
//   D5 self-loop transfer moves the record weight w, not the entry sum 2w (hit_leiden.rs:1681-1692)
//   D6 vertices added to a level by id-space growth get fresh ids (identity when nothing collides)
//   D7 the skip-aggregation path passes node_count = max(s_cur u s_pre)+1 upward (hit_leiden.rs:982-989)
//   D8 (flag OR_CONNECTED, parallel mode only) a singleton that is the target of another
//      singleton's merge proposal keeps its place, so every refined community stays
//      connected; the plain throughput mode applies all snapshot proposals (throughput.rs:278-287)
//   D9 the deterministic sweep skips self-loop entries in w(v, C) as the throughput shard
//      does (parallel_frontier.rs:164-166); the reference counts them (hit_leiden.rs:1119-1125),
//      which only matters on levels >= 1 (level-0 graphs carry no self-loops)
//   D10 the throughput movement of a level first relabels its community labels by rank
//      (order kept, so no tie decision changes; compact_labels): every label is then below
//      n and the empty-community ids n + v (parallel_frontier.rs:214) are always free.  The
//      reference's n + v can name a populated community once a level holds labels it did
//      not make (pulled down from above, fresh ids of grown vertices, earlier empty ids at a
//      smaller n); mode bit OR_REF_IDS skips the relabel, as the reference
//   D11 (OR_CONNECTED) the connected-component split links two vertices of a subcommunity
//      only through edges inside their community (f(u) = f(v)), the edges the spec's
//      subcommunity graph G_Psi keeps (hit_leiden_spec.md Alg. 2 lines 17-19, Alg. 3
//      lines 2-4); the reference links all same-subcommunity neighbours (hit_leiden.rs:
//      1233-1240, Psi not implemented), so a subcommunity whose vertex moved can straddle
//      two communities
//   D12 (OR_CONNECTED) def_update pulls every vertex of levels below the top, not only B^p /
//      R^p (hit_leiden.rs:1733-1793): f^p = f^{p+1} o s^p holds after every batch
#include "hl_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <queue>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#define HLO_SORT __gnu_parallel::sort
#else
#define HLO_SORT std::sort
#endif

namespace {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr double kEpsApply = 1e-12;     // in_memory.rs:110
constexpr double kEpsSelf = 1e-12;      // hit_leiden.rs:1687
constexpr double kEpsCompress = 1e-9;   // hit_leiden.rs:1711
constexpr double kEpsTFilter = 1e-9;    // hit_leiden.rs:1547,1631
constexpr int kMaxMoveRounds = 20;      // hit_leiden.rs:1077

// test hook: throughput moves to an empty-community id whose total was not
// zero when the round's moves were selected (a populated community joined as
// if it were empty)
u64 g_empty_collisions = 0;


struct Rec {
  u32 u, v;
  double w;
};

// Canonical CSR (in_memory.rs:20-73 with sorted rows; a self-loop is one entry).
struct Graph {
  u32 n = 0;
  std::vector<std::vector<u32>> col;
  std::vector<std::vector<double>> wt;
  std::vector<double> k;  // weighted degree: the row sum, self-loop counted twice
  double twice_m = 0.0;   // in_memory.rs:62 (x2), canonical pairwise-tree order
};

double row_degree(const Graph& g, u32 v) {
  double acc = 0.0;
  const auto& c = g.col[v];
  const auto& w = g.wt[v];
  for (size_t i = 0; i < c.size(); ++i) {
    acc += w[i];
    if (c[i] == v) acc += w[i];
  }
  return acc;
}

// canonical 2m: pairwise (halving) tree over x[0..n) zero-padded to a power of
// two -- at each level element i (a multiple of 2s) absorbs element i+s
double tree_sum(const std::vector<double>& x, u64 n) {
  if (n == 0) return 0.0;
  u64 N = 1;
  while (N < n) N <<= 1;
  std::vector<double> b(N, 0.0);
  std::copy(x.begin(), x.begin() + n, b.begin());
  for (u64 s = 1; s < N; s <<= 1)
    for (u64 i = 0; i + s < N; i += 2 * s) b[i] += b[i + s];
  return b[0];
}

void grow_graph(Graph& g, u32 n_new) {
  if (n_new <= g.n) return;
  g.col.resize(n_new);
  g.wt.resize(n_new);
  g.k.resize(n_new, 0.0);
  g.n = n_new;
}

// Returns index of neighbour x in row r or -1.
long find_entry(const Graph& g, u32 r, u32 x) {
  const auto& c = g.col[r];
  auto it = std::lower_bound(c.begin(), c.end(), x);
  if (it == c.end() || *it != x) return -1;
  return long(it - c.begin());
}

void set_entry(Graph& g, u32 r, u32 x, double w) {
  auto& c = g.col[r];
  auto it = std::lower_bound(c.begin(), c.end(), x);
  size_t pos = size_t(it - c.begin());
  if (it != c.end() && *it == x) {
    g.wt[r][pos] = w;
  } else {
    c.insert(it, x);
    g.wt[r].insert(g.wt[r].begin() + pos, w);
  }
}

void erase_entry(Graph& g, u32 r, u32 x) {
  long p = find_entry(g, r, x);
  if (p < 0) return;
  g.col[r].erase(g.col[r].begin() + p);
  g.wt[r].erase(g.wt[r].begin() + p);
}

void finish_degrees(Graph& g, const std::vector<u32>& touched) {
  for (u32 r : touched) g.k[r] = row_degree(g, r);
  g.twice_m = tree_sum(g.k, g.n);
}

// apply_delta / collect_canonical_edges / apply_delta_edges (in_memory.rs:103-138):
// per canonical key (min,max): existing weight (or 0) + alpha_1 + alpha_2 + ...
// in delta order; the key survives iff the sum is > 1e-12.
void apply_delta(Graph& g, const std::vector<Rec>& d, u32 n_new) {
  grow_graph(g, std::max(g.n, n_new));
  std::vector<std::pair<u64, u32>> keyed(d.size());
  for (size_t i = 0; i < d.size(); ++i) {
    u32 a = std::min(d[i].u, d[i].v), b = std::max(d[i].u, d[i].v);
    keyed[i] = {(u64(a) << 32) | b, u32(i)};
  }
  // (key, index) pairs are unique, so ordering them equals a stable sort by key
  HLO_SORT(keyed.begin(), keyed.end());
  std::vector<u32> touched;
  size_t i = 0;
  while (i < keyed.size()) {
    size_t j = i;
    u64 key = keyed[i].first;
    while (j < keyed.size() && keyed[j].first == key) ++j;
    u32 a = u32(key >> 32), b = u32(key & 0xffffffffu);
    long p = find_entry(g, a, b);
    double acc = p >= 0 ? g.wt[a][size_t(p)] : 0.0;
    for (size_t t = i; t < j; ++t) acc += d[keyed[t].second].w;
    if (acc > kEpsApply) {
      set_entry(g, a, b, acc);
      if (a != b) set_entry(g, b, a, acc);
    } else if (p >= 0) {
      erase_entry(g, a, b);
      if (a != b) erase_entry(g, b, a);
    }
    touched.push_back(a);
    if (a != b) touched.push_back(b);
    i = j;
  }
  std::sort(touched.begin(), touched.end());
  touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
  finish_degrees(g, touched);
}

// apply_delta as written (in_memory.rs:103-131): every edge of the graph into
// a hash map keyed by (min, max) (collect_canonical_edges), the delta added
// per key in delta order (apply_delta_edges), the graph rebuilt from the keys
// above 1e-12 (InMemoryGraph::from) and every degree recomputed.  The map adds
// each key's weights in the same order as apply_delta, and the rows are
// sorted again, so the result is apply_delta's; only the cost is the
// reference's, O(n + m) per batch and level (the CPU baseline's variant (i)).
void apply_delta_rebuild(Graph& g, const std::vector<Rec>& d, u32 n_new) {
  grow_graph(g, std::max(g.n, n_new));
  std::unordered_map<u64, double> e;
  e.reserve(size_t(g.twice_m > 0 ? g.n : 0) + d.size());
  for (u32 a = 0; a < g.n; ++a)
    for (size_t j = 0; j < g.col[a].size(); ++j) {
      const u32 b = g.col[a][j];
      if (a <= b) e[(u64(a) << 32) | b] = g.wt[a][j];
    }
  for (const Rec& r : d) {
    const u32 a = std::min(r.u, r.v), b = std::max(r.u, r.v);
    e[(u64(a) << 32) | b] += r.w;  // a new key starts from 0.0
  }
  for (u32 v = 0; v < g.n; ++v) {
    g.col[v].clear();
    g.wt[v].clear();
  }
  for (const auto& kv : e) {
    if (!(kv.second > kEpsApply)) continue;
    const u32 a = u32(kv.first >> 32), b = u32(kv.first & 0xffffffffu);
    g.col[a].push_back(b);
    g.wt[a].push_back(kv.second);
    if (a != b) {
      g.col[b].push_back(a);
      g.wt[b].push_back(kv.second);
    }
  }
  std::vector<std::pair<u32, double>> row;
  for (u32 v = 0; v < g.n; ++v) {
    row.resize(g.col[v].size());
    for (size_t j = 0; j < row.size(); ++j) row[j] = {g.col[v][j], g.wt[v][j]};
    std::sort(row.begin(), row.end(), [](const std::pair<u32, double>& x, const std::pair<u32, double>& y) {
      return x.first < y.first;
    });
    for (size_t j = 0; j < row.size(); ++j) {
      g.col[v][j] = row[j].first;
      g.wt[v][j] = row[j].second;
    }
    g.k[v] = row_degree(g, v);
  }
  g.twice_m = tree_sum(g.k, g.n);
}

// The reference's O(n) scans (the CPU baseline's variant (i)): the members of
// subcommunity sc in ascending order, and the degree total of label c summed
// in vertex order -- the values the member index and the totals hold.
std::vector<u32> members_scan(const std::vector<u32>& s, u32 n, u32 sc) {
  std::vector<u32> m;
  for (u32 i = 0; i < n; ++i)
    if (s[i] == sc) m.push_back(i);
  return m;
}
double total_scan(const Graph& g, const std::vector<u32>& lab, u32 c) {
  double acc = 0.0;
  for (u32 i = 0; i < g.n; ++i)
    if (lab[i] == c) acc += g.k[i];
  return acc;
}

Graph graph_from_edges(u32 n, u64 m, const u32* eu, const u32* ev, const double* ew) {
  Graph g;
  grow_graph(g, n);
  std::vector<Rec> d(m);
  for (u64 i = 0; i < m; ++i) d[i] = {eu[i], ev[i], ew ? ew[i] : 1.0};
  apply_delta(g, d, n);
  return g;
}

// Community-degree totals: per label, sequential sum of k_v in vertex order,
// array sized max(n, max label + 1) (hit_leiden.rs:1059-1066 with D1).
std::vector<double> label_totals(const Graph& g, const std::vector<u32>& lab, size_t extra = 0) {
  u64 size = g.n;
  for (u32 v = 0; v < g.n; ++v) size = std::max<u64>(size, u64(lab[v]) + 1);
  std::vector<double> tot(size + extra, 0.0);
  for (u32 v = 0; v < g.n; ++v) tot[lab[v]] += g.k[v];
  tot.resize(size);
  return tot;
}

// ---------------------------------------------------------------------------
// Movement (Algorithm 2): inc_movement hit_leiden.rs:1024-1192 (deterministic),
// inc_movement_parallel / execute_shard / select_decoupled_moves
// (throughput.rs:80-236, parallel_frontier.rs:133-239) for the throughput mode.
// ---------------------------------------------------------------------------
struct MoveCand {
  u32 node, from, to;
  double deg, gain;
};

inline double move_gain(double w_cand, double w_cur, double t, double gamma, double k, double sc, double scand) {
  // hit_leiden.rs:1135-1140 -- same operation order, no contraction
  return (w_cand - w_cur) / t + gamma * k * (sc - k - scand) / (t * t);
}
inline double empty_gain(double w_cur, double t, double gamma, double k, double sc) {
  // hit_leiden.rs:1150-1154
  return -w_cur / t + gamma * k * (sc - k) / (t * t);
}

// select_decoupled_moves (throughput.rs:204-236): sort by (gain desc, node asc),
// greedy accept unless from is an acceptor or to is an emitter, one move per node.
std::vector<MoveCand> select_decoupled(std::vector<MoveCand> c) {
  std::stable_sort(c.begin(), c.end(), [](const MoveCand& a, const MoveCand& b) {
    if (a.gain != b.gain) return a.gain > b.gain;
    return a.node < b.node;
  });
  std::vector<u32> emit, acc, moved;
  auto has = [](const std::vector<u32>& s, u32 x) { return std::binary_search(s.begin(), s.end(), x); };
  auto add = [](std::vector<u32>& s, u32 x) {
    auto it = std::lower_bound(s.begin(), s.end(), x);
    if (it == s.end() || *it != x) s.insert(it, x);
  };
  std::vector<MoveCand> sel;
  // small-set vectors are fine for tests; large runs use the flag arrays below
  if (c.size() > 4096) {
    u64 maxid = 0, maxnode = 0;
    for (auto& x : c) {
      maxid = std::max<u64>(maxid, std::max(x.from, x.to));
      maxnode = std::max<u64>(maxnode, x.node);
    }
    std::vector<uint8_t> fe(maxid + 1, 0), fa(maxid + 1, 0), fm(maxnode + 1, 0);
    for (auto& x : c) {
      if (x.gain <= 0.0) continue;
      if (fm[x.node]) continue;
      if (fa[x.from] || fe[x.to]) continue;
      sel.push_back(x);
      fm[x.node] = 1;
      fe[x.from] = 1;
      fa[x.to] = 1;
    }
    return sel;
  }
  for (auto& x : c) {
    if (x.gain <= 0.0) continue;
    if (has(moved, x.node)) continue;
    if (has(acc, x.from) || has(emit, x.to)) continue;
    sel.push_back(x);
    add(moved, x.node);
    add(emit, x.from);
    add(acc, x.to);
  }
  return sel;
}

struct MoveResult {
  std::vector<uint8_t> A0, B, K;
  u32 iters = 0;
};

// execute_shard's per-vertex body (parallel_frontier.rs:155-236): self-loops
// skipped, candidates in ascending id, empty community id = ebase + v (callers
// pass ebase = n: the reference's n + v).
bool propose_thr(const Graph& g, u32 v, const std::vector<u32>& f, const std::vector<double>& tot, double t,
                 double gamma, u64 ebase, std::vector<std::pair<u32, double>>& buf, MoveCand& out) {
  const u32 cur = f[v];
  const double k = g.k[v];
  buf.clear();
  double w_cur = 0.0;
  const auto& c = g.col[v];
  const auto& w = g.wt[v];
  for (size_t i = 0; i < c.size(); ++i) {
    if (c[i] == v) continue;
    buf.push_back({f[c[i]], w[i]});
  }
  // per-community sums in row order
  std::stable_sort(buf.begin(), buf.end(),
                   [](const std::pair<u32, double>& a, const std::pair<u32, double>& b) { return a.first < b.first; });
  for (auto& e : buf)
    if (e.first == cur) w_cur += e.second;
  const double sc = tot[cur];
  u32 best = cur;
  double best_gain = 0.0;
  size_t i = 0;
  while (i < buf.size()) {
    u32 cc = buf[i].first;
    double acc = 0.0;
    while (i < buf.size() && buf[i].first == cc) acc += buf[i++].second;
    if (cc == cur) continue;
    double gain = move_gain(acc, w_cur, t, gamma, k, sc, tot[cc]);
    if (gain > best_gain) {
      best_gain = gain;
      best = cc;
    }
  }
  double eg = empty_gain(w_cur, t, gamma, k, sc);
  if (eg > best_gain) {
    best_gain = eg;
    best = u32(ebase + v);
  }
  if (best_gain <= 0.0) return false;
  out = {v, cur, best, k, best_gain};
  return true;
}

MoveResult inc_movement(const Graph& g, const std::vector<Rec>& dG, std::vector<u32>& f, const std::vector<u32>& s,
                        double gamma, int mode, u64 ebase) {
  const u32 n = g.n;
  MoveResult r;
  std::vector<uint8_t> A(n, 0);
  r.B.assign(n, 0);
  r.K.assign(n, 0);
  // seeding, hit_leiden.rs:1037-1052
  for (const Rec& e : dG) {
    double alpha = e.w;
    if (alpha > 0.0 && f[e.u] != f[e.v]) A[e.u] = A[e.v] = 1;
    if (alpha < 0.0 && f[e.u] == f[e.v]) A[e.u] = A[e.v] = 1;
    if (s[e.u] == s[e.v]) r.K[e.u] = r.K[e.v] = 1;
  }
  if (dG.empty()) std::fill(A.begin(), A.end(), 1);  // hit_leiden.rs:1054-1057
  r.A0 = A;
  const double t = g.twice_m;
  std::vector<double> tot = label_totals(g, f);

  if (mode & OR_THR) {
    std::vector<u32> active;
    for (u32 v = 0; v < n; ++v)
      if (A[v]) active.push_back(v);
    u32 it = 0;
    std::vector<uint8_t> next(n, 0);
    while (!active.empty() && it < u32(kMaxMoveRounds)) {
      std::vector<MoveCand> props(active.size());
      std::vector<uint8_t> has(active.size(), 0);
#pragma omp parallel
      {
        std::vector<std::pair<u32, double>> buf;
#pragma omp for schedule(dynamic, 256)
        for (long i = 0; i < long(active.size()); ++i) has[i] = propose_thr(g, active[i], f, tot, t, gamma, ebase, buf, props[i]);
      }
      std::vector<MoveCand> cand;
      for (size_t i = 0; i < active.size(); ++i)
        if (has[i]) cand.push_back(props[i]);
      std::vector<MoveCand> sel = select_decoupled(std::move(cand));
      ++it;
      u64 max_to = 0;
      for (auto& m : sel) max_to = std::max<u64>(max_to, m.to);
      if (!sel.empty() && max_to >= tot.size()) tot.resize(max_to + 1, 0.0);
      double total_gain = 0.0;
      for (auto& m : sel) total_gain += m.gain;
      if (total_gain <= 0.0) break;  // throughput.rs:177-180
      for (auto& m : sel)
        if (u64(m.to) == ebase + m.node && m.to != m.from && tot[m.to] != 0.0) ++g_empty_collisions;
      for (auto& m : sel) {
        f[m.node] = m.to;
        tot[m.from] -= m.deg;
        tot[m.to] += m.deg;
        r.B[m.node] = 1;
      }
      std::fill(next.begin(), next.end(), 0);
      for (auto& m : sel) {
        const auto& c = g.col[m.node];
        for (u32 u : c) {
          if (f[u] != m.to) next[u] = 1;
          if (s[m.node] == s[u]) r.K[m.node] = r.K[u] = 1;
        }
      }
      active.clear();
      for (u32 v = 0; v < n; ++v)
        if (next[v]) active.push_back(v);
    }
    r.iters = it;
    return r;
  }

  // deterministic: pop the lowest-index active vertex (hit_leiden.rs:1108-1188)
  u64 next_empty = tot.size();
  std::priority_queue<u32, std::vector<u32>, std::greater<u32>> pq;
  for (u32 v = 0; v < n; ++v)
    if (A[v]) pq.push(v);
  std::vector<std::pair<u32, double>> buf;
  while (!pq.empty()) {
    u32 v = pq.top();
    pq.pop();
    if (!A[v]) continue;
    A[v] = 0;
    const u32 cur = f[v];
    const double k = g.k[v];
    buf.clear();
    const auto& c = g.col[v];
    const auto& w = g.wt[v];
    for (size_t i = 0; i < c.size(); ++i) {
      if (c[i] == v) continue;  // D9: self-loops skipped as in parallel_frontier.rs:164-166
      buf.push_back({f[c[i]], w[i]});
    }
    std::stable_sort(buf.begin(), buf.end(),
                     [](const std::pair<u32, double>& a, const std::pair<u32, double>& b) { return a.first < b.first; });
    double w_cur = 0.0;
    for (auto& e : buf)
      if (e.first == cur) w_cur += e.second;
    u32 best = cur;
    double best_gain = 0.0;
    const double sc = tot[cur];
    size_t i = 0;
    while (i < buf.size()) {
      u32 cc = buf[i].first;
      double acc = 0.0;
      while (i < buf.size() && buf[i].first == cc) acc += buf[i++].second;
      if (cc == cur) continue;
      double gain = move_gain(acc, w_cur, t, gamma, k, sc, tot[cc]);
      if (gain > best_gain) {
        best_gain = gain;
        best = cc;
      }
    }
    double eg = empty_gain(w_cur, t, gamma, k, sc);
    u64 best64 = best;
    if (eg > best_gain) {
      best_gain = eg;
      best64 = next_empty;
    }
    if (best_gain <= 0.0) continue;
    if (best64 >= tot.size()) {
      tot.resize(best64 + 1, 0.0);
      next_empty = best64 + 1;
    }
    best = u32(best64);
    f[v] = best;
    r.B[v] = 1;
    tot[cur] -= k;
    tot[best] += k;
    for (size_t j = 0; j < c.size(); ++j) {
      u32 u = c[j];
      if (f[u] != best && !A[u]) {
        A[u] = 1;
        pq.push(u);
      }
      if (s[v] == s[u]) r.K[v] = r.K[u] = 1;
    }
  }
  r.iters = 1;
  return r;
}

// ---------------------------------------------------------------------------
// Refinement (Algorithm 3): inc_refinement hit_leiden.rs:1265-1416, singleton
// merge + T-filter :1419-1548, blocked set :1554-1637, throughput proposals
// throughput.rs:239-359.
// ---------------------------------------------------------------------------
struct SubIndex {  // members of each subcommunity, ascending vertex
  std::vector<u64> off;
  std::vector<u32> mem;
};

SubIndex build_members(const std::vector<u32>& s, u32 n, u64 size) {
  SubIndex ix;
  ix.off.assign(size + 1, 0);
  for (u32 v = 0; v < n; ++v) ix.off[s[v] + 1]++;
  for (u64 i = 0; i < size; ++i) ix.off[i + 1] += ix.off[i];
  ix.mem.resize(n);
  std::vector<u64> pos(ix.off.begin(), ix.off.end() - 1);
  for (u32 v = 0; v < n; ++v) ix.mem[pos[s[v]]++] = v;
  return ix;
}

// w(S, C) = sum over members i of S (ascending) and row entries of i (row order,
// self-loop counted twice) with f(nbr) = C.  hit_leiden.rs:1535-1544 / 1617-1629.
double weight_to_comm(const Graph& g, const std::vector<u32>& members, const std::vector<u32>& f, u32 C) {
  double acc = 0.0;
  for (u32 i : members) {
    const auto& c = g.col[i];
    const auto& w = g.wt[i];
    for (size_t j = 0; j < c.size(); ++j) {
      if (f[c[j]] == C) {
        acc += w[j];
        if (c[j] == i) acc += w[j];
      }
    }
  }
  return acc;
}

inline double remove_gain(double w_s, double t, double gamma, double d_s, double d_parent) {
  // hit_leiden.rs:1546-1547, :1630-1631
  return -w_s / t + gamma * d_s * (d_parent - d_s) / (t * t);
}

// connected-component split of the affected subcommunities (hit_leiden.rs:1276-1331)
void cc_split(const Graph& g, std::vector<u32>& s, const std::vector<uint8_t>& K, std::vector<uint8_t>& R,
              const std::vector<u32>* f) {
  const u32 n = g.n;
  if (n == 0) return;
  u64 smax = 0;
  for (u32 v = 0; v < n; ++v) smax = std::max<u64>(smax, s[v]);
  std::vector<uint8_t> aff(smax + 1, 0);
  bool any = false;
  for (u32 v = 0; v < n; ++v)
    if (K[v]) aff[s[v]] = 1, any = true;
  if (!any) return;
  u64 next_id = smax + 1;
  SubIndex ix = build_members(s, n, smax + 1);
  std::vector<uint8_t> visited(n, 0);
  std::vector<u32> queue;
  for (u64 sc = 0; sc <= smax; ++sc) {
    if (!aff[sc]) continue;
    std::vector<std::vector<u32>> comps;
    for (u64 p = ix.off[sc]; p < ix.off[sc + 1]; ++p) {
      u32 start = ix.mem[p];
      if (visited[start]) continue;
      queue.clear();
      queue.push_back(start);
      visited[start] = 1;
      size_t h = 0;
      while (h < queue.size()) {
        u32 x = queue[h++];
        for (u32 y : g.col[x])
          if (s[y] == u32(sc) && !visited[y] && (!f || (*f)[y] == (*f)[x])) {  // D11
            visited[y] = 1;
            queue.push_back(y);
          }
      }
      comps.push_back(queue);
    }
    for (u64 p = ix.off[sc]; p < ix.off[sc + 1]; ++p) visited[ix.mem[p]] = 0;
    if (comps.size() <= 1) continue;
    size_t keep = 0;
    for (size_t c = 0; c < comps.size(); ++c)
      if (comps[c].size() >= comps[keep].size()) keep = c;  // last largest (max_by_key)
    for (size_t c = 0; c < comps.size(); ++c) {
      if (c == keep) continue;
      u32 nid = u32(next_id++);
      for (u32 v : comps[c]) {
        s[v] = nid;
        R[v] = 1;
      }
    }
  }
}

std::vector<uint8_t> inc_refinement(const Graph& g, const std::vector<u32>& f, std::vector<u32>& s,
                                    const std::vector<uint8_t>& K, double gamma, int mode) {
  const u32 n = g.n;
  std::vector<uint8_t> R(n, 0);
  cc_split(g, s, K, R, (mode & OR_CONNECTED) ? &f : nullptr);
  bool is_initial = true;
  for (u32 v = 0; v < n; ++v)
    if (s[v] != v) {
      is_initial = false;
      break;
    }
  if (is_initial) std::fill(R.begin(), R.end(), 1);  // hit_leiden.rs:1333-1339
  if (n == 0) return R;
  u64 smax = 0;
  for (u32 v = 0; v < n; ++v) smax = std::max<u64>(smax, s[v]);
  std::vector<u64> sizes(smax + 1, 0);
  for (u32 v = 0; v < n; ++v) sizes[s[v]]++;
  std::vector<double> dS(smax + 1, 0.0);  // vertex-order sums (hit_leiden.rs:1348-1357)
  for (u32 v = 0; v < n; ++v) dS[s[v]] += g.k[v];
  const double t = g.twice_m;
  std::vector<u32> Rs;
  for (u32 v = 0; v < n; ++v)
    if (R[v]) Rs.push_back(v);
  std::stable_sort(Rs.begin(), Rs.end(), [&](u32 a, u32 b) { return g.k[a] < g.k[b]; });
  std::vector<double> dparent = label_totals(g, f);  // per community, vertex order

  if (mode & OR_THR) {
    // compute_blocked_subcommunities (hit_leiden.rs:1554-1637)
    std::vector<uint8_t> is_cand(smax + 1, 0);
    for (u32 v : Rs) {
      if (sizes[s[v]] != 1) continue;
      for (u32 u : g.col[v])
        if (f[u] == f[v] && s[u] != s[v]) is_cand[s[u]] = 1;
    }
    SubIndex ix = build_members(s, n, smax + 1);
    std::vector<uint8_t> blocked(smax + 1, 0);
    std::vector<u32> cands;
    for (u64 sc = 0; sc <= smax; ++sc)
      if (is_cand[sc]) cands.push_back(u32(sc));
#pragma omp parallel for schedule(dynamic, 64)
    for (long ci = 0; ci < long(cands.size()); ++ci) {
      u32 sc = cands[ci];
      double d_s = dS[sc];
      if (!(d_s > 0.0)) continue;
      u64 b = ix.off[sc], e = ix.off[sc + 1];
      u32 parent = b < e ? f[ix.mem[b]] : 0;
      double d_par = parent < dparent.size() ? dparent[parent] : 0.0;
      std::vector<u32> mem(ix.mem.begin() + b, ix.mem.begin() + e);
      if (mode & OR_AS_WRITTEN) {  // hit_leiden.rs:1604-1629: the parent and the members by O(n) scans
        mem = members_scan(s, n, sc);
        parent = mem.empty() ? 0 : f[mem[0]];
      }
      double w_s = weight_to_comm(g, mem, f, parent);
      if (remove_gain(w_s, t, gamma, d_s, d_par) > kEpsTFilter) blocked[sc] = 1;
    }
    // try_refinement_move (throughput.rs:291-359) against the snapshot
    std::vector<u32> prop(Rs.size(), UINT32_MAX);
#pragma omp parallel
    {
      std::vector<std::pair<u32, double>> buf;
#pragma omp for schedule(dynamic, 256)
      for (long i = 0; i < long(Rs.size()); ++i) {
        u32 v = Rs[i];
        u32 sc = s[v];
        if (sizes[sc] != 1) continue;
        const u32 C = f[v];
        const double k = g.k[v];
        buf.clear();
        const auto& c = g.col[v];
        const auto& w = g.wt[v];
        for (size_t j = 0; j < c.size(); ++j) {
          if (f[c[j]] != C) continue;
          buf.push_back({s[c[j]], w[j]});
          if (c[j] == v) buf.push_back({s[c[j]], w[j]});
        }
        std::stable_sort(buf.begin(), buf.end(), [](const std::pair<u32, double>& a,
                                                    const std::pair<u32, double>& b) { return a.first < b.first; });
        double w_cur = 0.0;
        for (auto& e : buf)
          if (e.first == sc) w_cur += e.second;
        u32 best = sc;
        double best_gain = 0.0;
        size_t p = 0;
        while (p < buf.size()) {
          u32 cs = buf[p].first;
          double acc = 0.0;
          while (p < buf.size() && buf[p].first == cs) acc += buf[p++].second;
          if (cs == sc || blocked[cs]) continue;
          double gain = move_gain(acc, w_cur, t, gamma, k, dS[sc], dS[cs]);
          if (gain > best_gain) {
            best_gain = gain;
            best = cs;
          }
        }
        if (best_gain > 0.0) prop[i] = best;
      }
    }
    if (mode & OR_CONNECTED) {
      std::vector<uint8_t> target(smax + 1, 0);
      for (size_t i = 0; i < Rs.size(); ++i)
        if (prop[i] != UINT32_MAX) target[prop[i]] = 1;
      for (size_t i = 0; i < Rs.size(); ++i)
        if (prop[i] != UINT32_MAX && target[s[Rs[i]]]) prop[i] = UINT32_MAX;
    }
    for (size_t i = 0; i < Rs.size(); ++i)
      if (prop[i] != UINT32_MAX) s[Rs[i]] = prop[i];
    return R;
  }

  // deterministic singleton merges in (k asc, id asc) order
  SubIndex ix = build_members(s, n, smax + 1);
  std::vector<std::vector<u32>> members(smax + 1);
  for (u64 sc = 0; sc <= smax; ++sc) members[sc].assign(ix.mem.begin() + ix.off[sc], ix.mem.begin() + ix.off[sc + 1]);
  std::vector<std::pair<u32, double>> buf;
  for (u32 v : Rs) {
    u32 sc = s[v];
    if (sizes[sc] != 1) continue;
    const u32 C = f[v];
    const double k = g.k[v];
    buf.clear();
    const auto& c = g.col[v];
    const auto& w = g.wt[v];
    for (size_t j = 0; j < c.size(); ++j) {
      if (f[c[j]] != C) continue;
      buf.push_back({s[c[j]], w[j]});
      if (c[j] == v) buf.push_back({s[c[j]], w[j]});
    }
    std::stable_sort(buf.begin(), buf.end(),
                     [](const std::pair<u32, double>& a, const std::pair<u32, double>& b) { return a.first < b.first; });
    double w_cur = 0.0;
    for (auto& e : buf)
      if (e.first == sc) w_cur += e.second;
    u32 best = sc;
    double best_gain = 0.0;
    // hit_leiden.rs:1454-1458 sums the parent's degrees by an O(n) scan
    const double d_par = (mode & OR_AS_WRITTEN) ? total_scan(g, f, C) : (C < dparent.size() ? dparent[C] : 0.0);
    size_t p = 0;
    while (p < buf.size()) {
      u32 cs = buf[p].first;
      double acc = 0.0;
      while (p < buf.size() && buf[p].first == cs) acc += buf[p++].second;
      if (cs == sc) continue;
      double d_s = dS[cs];
      if (d_s > 0.0) {  // should_skip_candidate_t_filter
        double w_s = (mode & OR_AS_WRITTEN) ? weight_to_comm(g, members_scan(s, n, cs), f, C)  // :1535-1544
                                            : weight_to_comm(g, members[cs], f, C);
        if (remove_gain(w_s, t, gamma, d_s, d_par) > kEpsTFilter) continue;
      }
      double gain = move_gain(acc, w_cur, t, gamma, k, dS[sc], dS[cs]);
      if (gain > best_gain) {
        best_gain = gain;
        best = cs;
      }
    }
    if (best_gain > 0.0) {
      s[v] = best;
      sizes[sc] -= 1;
      sizes[best] += 1;
      dS[sc] -= k;
      dS[best] += k;
      auto& ms = members[sc];
      ms.erase(std::find(ms.begin(), ms.end(), v));
      auto& mb = members[best];
      mb.insert(std::lower_bound(mb.begin(), mb.end(), v), v);
    }
  }
  return R;
}

// ---------------------------------------------------------------------------
// Aggregation (Algorithm 4): inc_aggregation hit_leiden.rs:1639-1731 (+D4, D5).
// ---------------------------------------------------------------------------
std::vector<Rec> inc_aggregation(const Graph& g, const std::vector<Rec>& dG, std::vector<u32>& s_pre,
                                 const std::vector<u32>& s_cur, const std::vector<uint8_t>& R, u32& next_n) {
  std::vector<Rec> H;
  for (const Rec& e : dG) H.push_back({s_pre[e.u], s_pre[e.v], e.w});
  for (u32 v = 0; v < g.n; ++v) {
    if (!R[v]) continue;
    const auto& c = g.col[v];
    const auto& w = g.wt[v];
    double self = 0.0;
    for (size_t j = 0; j < c.size(); ++j) {
      u32 u = c[j];
      if (u == v) {
        self = w[j];
        continue;
      }
      if (!R[u] || v < u) {
        H.push_back({s_pre[v], s_pre[u], -w[j]});
        H.push_back({s_cur[v], s_cur[u], w[j]});
      }
    }
    if (std::fabs(self) > kEpsSelf) {
      H.push_back({s_pre[v], s_pre[v], -self});
      H.push_back({s_cur[v], s_cur[v], self});
    }
  }
  u64 mx = 0;
  bool any = false;
  for (u32 v = 0; v < g.n; ++v) {
    mx = std::max<u64>(mx, std::max(s_cur[v], s_pre[v]));
    any = true;
  }
  next_n = any ? u32(mx + 1) : 1;
  for (u32 v = 0; v < g.n; ++v)
    if (R[v]) s_pre[v] = s_cur[v];
  // Compress: per canonical key in emission order, keep |w| > 1e-9, sorted output
  std::vector<std::pair<u64, u32>> keyed(H.size());
  for (size_t i = 0; i < H.size(); ++i) {
    u32 a = std::min(H[i].u, H[i].v), b = std::max(H[i].u, H[i].v);
    keyed[i] = {(u64(a) << 32) | b, u32(i)};
  }
  // (key, index) pairs are unique, so ordering them equals a stable sort by key
  HLO_SORT(keyed.begin(), keyed.end());
  std::vector<Rec> out;
  size_t i = 0;
  while (i < keyed.size()) {
    size_t j = i;
    double acc = 0.0;
    while (j < keyed.size() && keyed[j].first == keyed[i].first) acc += H[keyed[j++].second].w;
    if (std::fabs(acc) > kEpsCompress)
      out.push_back({u32(keyed[i].first >> 32), u32(keyed[i].first & 0xffffffffu), acc});
    i = j;
  }
  return out;
}

// def_update (hit_leiden.rs:1733-1793), per-level sizes.
void def_update(std::vector<std::vector<u32>*>& maps, const std::vector<const std::vector<u32>*>& sub,
                std::vector<std::vector<uint8_t>>& B, bool all = false) {
  int P = int(maps.size());
  for (int p = P - 1; p >= 0; --p) {
    if (p < P - 1) {
      auto& fp = *maps[p];
      const auto& fq = *maps[p + 1];
      const auto& sp = *sub[p];
      for (size_t v = 0; v < B[p].size(); ++v)
        if (all || B[p][v]) fp[v] = fq[sp[v]];  // D12 with all
      if (all) continue;  // every vertex pulled: the B propagation changes nothing
    }
    if (p > 0) {
      const auto& sq = *sub[p - 1];
      for (size_t u = 0; u < B[p - 1].size(); ++u) {
        u32 x = sq[u];
        if (x < B[p].size() && B[p][x]) B[p - 1][u] = 1;
      }
    }
  }
}

double modularity(const Graph& g, const u32* lab) {
  const double t = g.twice_m;
  if (t == 0.0) return 0.0;
  double intra = 0.0;
  for (u32 v = 0; v < g.n; ++v) {
    const auto& c = g.col[v];
    const auto& w = g.wt[v];
    for (size_t j = 0; j < c.size(); ++j)
      if (lab[c[j]] == lab[v]) intra += (c[j] == v) ? 2.0 * w[j] : w[j];
  }
  u64 size = g.n;
  for (u32 v = 0; v < g.n; ++v) size = std::max<u64>(size, u64(lab[v]) + 1);
  std::vector<double> sig(size, 0.0);
  for (u32 v = 0; v < g.n; ++v) sig[lab[v]] += g.k[v];
  double expected = 0.0;
  for (double x : sig) expected += x * x / t;
  return (intra - expected) / t;
}

// ---------------------------------------------------------------------------
// State + Algorithm 6 (hit_leiden.rs:936-1018) with P levels.
// ---------------------------------------------------------------------------
struct Level {
  Graph G;
  std::vector<u32> f, g, s_pre, s_cur;
  std::vector<uint8_t> A0, B, K, R;
  std::vector<Rec> dG;
  u32 iters = 0;
};

// D10: rank relabel of a level's community labels (ascending label order kept)
void compact_labels(std::vector<u32>& f) {
  std::vector<u32> vals(f);
  std::sort(vals.begin(), vals.end());
  vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
  for (auto& x : f) x = u32(std::lower_bound(vals.begin(), vals.end(), x) - vals.begin());
}

struct State {
  int P = 1;
  double gamma = 1.0;
  int mode = OR_DET;
  std::vector<Level> L;
  u32 n0 = 0;
};

u64 vmax_or(const std::vector<u32>& x, u64 dflt) {
  if (x.empty()) return dflt;
  return *std::max_element(x.begin(), x.end());
}

// D6: new vertices [n_old, n_new) get fresh labels.
void grow_level(Level& lv, u32 n_new, const Level* up) {
  u32 n_old = lv.G.n;
  if (n_new <= n_old) return;
  u64 base_f = std::max<u64>(n_old, lv.f.empty() ? 0 : vmax_or(lv.f, 0) + 1);
  u64 base_g = std::max<u64>(n_old, lv.g.empty() ? 0 : vmax_or(lv.g, 0) + 1);
  u64 ms = 0;
  bool have = !lv.s_cur.empty();
  if (have) ms = std::max(vmax_or(lv.s_cur, 0), vmax_or(lv.s_pre, 0)) + 1;
  u64 base_s = std::max<u64>(n_old, ms);
  if (up) base_s = std::max<u64>(base_s, up->G.n);
  for (u32 v = n_old; v < n_new; ++v) {
    lv.f.push_back(u32(base_f + (v - n_old)));
    lv.g.push_back(u32(base_g + (v - n_old)));
    lv.s_pre.push_back(u32(base_s + (v - n_old)));
    lv.s_cur.push_back(u32(base_s + (v - n_old)));
  }
  grow_graph(lv.G, n_new);
}

u32 hit_leiden(State& st, const std::vector<Rec>& dG0) {
  std::vector<Rec> cur = dG0;
  u32 cur_n = st.n0;
  u32 total = 0;
  const int P = st.P;
  for (int p = 0; p < P; ++p) {
    Level& lv = st.L[p];
    grow_level(lv, std::max(lv.G.n, cur_n), p + 1 < P ? &st.L[p + 1] : nullptr);
    if (st.mode & OR_AS_WRITTEN)
      apply_delta_rebuild(lv.G, cur, lv.G.n);
    else
      apply_delta(lv.G, cur, lv.G.n);
    lv.dG = cur;
    if ((st.mode & OR_THR) && !(st.mode & OR_REF_IDS)) compact_labels(lv.f);  // D10
    MoveResult mr = inc_movement(lv.G, cur, lv.f, lv.s_cur, st.gamma, st.mode, lv.G.n);
    lv.A0 = std::move(mr.A0);
    lv.B = std::move(mr.B);
    lv.K = std::move(mr.K);
    lv.iters = mr.iters;
    total += mr.iters;
    lv.R = inc_refinement(lv.G, lv.f, lv.s_cur, lv.K, st.gamma, st.mode);
    if (p < P - 1) {
      bool anyR = std::any_of(lv.R.begin(), lv.R.end(), [](uint8_t x) { return x != 0; });
      if (cur.empty() && !anyR) {  // should_skip_aggregation (hit_leiden.rs:1020-1022), D7
        u64 mx = 0;
        for (u32 v = 0; v < lv.G.n; ++v) mx = std::max<u64>(mx, std::max(lv.s_cur[v], lv.s_pre[v]));
        cur.clear();
        cur_n = lv.G.n ? u32(mx + 1) : 0;
        continue;
      }
      u32 nn = 0;
      cur = inc_aggregation(lv.G, cur, lv.s_pre, lv.s_cur, lv.R, nn);
      cur_n = nn;
    }
  }
  std::vector<std::vector<uint8_t>> B(P), R(P);
  std::vector<std::vector<u32>*> fm(P), gm(P);
  std::vector<const std::vector<u32>*> sm(P);
  for (int p = 0; p < P; ++p) {
    B[p] = st.L[p].B;
    R[p] = st.L[p].R;
    fm[p] = &st.L[p].f;
    gm[p] = &st.L[p].g;
    sm[p] = &st.L[p].s_cur;
  }
  const bool all = (st.mode & OR_CONNECTED) != 0;  // D12
  def_update(fm, sm, B, all);
  def_update(gm, sm, R, all);
  return total;
}

void reset_state(State& st, u32 n) {
  st.L.assign(st.P, Level());
  st.n0 = n;
  grow_level(st.L[0], n, st.P > 1 ? &st.L[1] : nullptr);
}

bool valid_edges(u32 n, u64 m, const u32* u, const u32* v, const double* w) {
  for (u64 i = 0; i < m; ++i) {
    if (u[i] >= n || v[i] >= n) return false;
    if (w && !std::isfinite(w[i])) return false;
  }
  return true;
}

std::vector<uint8_t>& level_set(Level& lv, int which) {
  switch (which) {
    case OR_SET_A: return lv.A0;
    case OR_SET_B: return lv.B;
    case OR_SET_K: return lv.K;
    default: return lv.R;
  }
}
std::vector<u32>& level_map(Level& lv, int which) {
  switch (which) {
    case OR_MAP_F: return lv.f;
    case OR_MAP_G: return lv.g;
    case OR_MAP_S_CUR: return lv.s_cur;
    default: return lv.s_pre;
  }
}

}  // namespace

extern "C" {

void* or_create(int levels, double gamma, int mode, int threads) {
  if (levels < 1) return nullptr;
  State* st = new State();
  st->P = levels;
  st->gamma = gamma;
  st->mode = mode;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  reset_state(*st, 0);
  return st;
}

void or_destroy(void* h) { delete static_cast<State*>(h); }

int or_load_graph(void* h, uint32_t n, uint64_t m, const uint32_t* u, const uint32_t* v, const double* w) {
  State& st = *static_cast<State*>(h);
  if (!valid_edges(n, m, u, v, w)) return -1;
  reset_state(st, n);
  std::vector<Rec> d(m);
  for (u64 i = 0; i < m; ++i) d[i] = {u[i], v[i], w ? w[i] : 1.0};
  apply_delta(st.L[0].G, d, n);
  return 0;
}

int or_load(void* h, uint32_t n, uint64_t m, const uint32_t* u, const uint32_t* v, const double* w) {
  State& st = *static_cast<State*>(h);
  if (!valid_edges(n, m, u, v, w)) return -1;
  reset_state(st, n);
  std::vector<Rec> d(m);
  for (u64 i = 0; i < m; ++i) d[i] = {u[i], v[i], w ? w[i] : 1.0};
  return int(hit_leiden(st, d));
}

int64_t or_apply(void* h, uint64_t k, const uint32_t* u, const uint32_t* v, const double* a) {
  State& st = *static_cast<State*>(h);
  if (!valid_edges(st.n0, k, u, v, a)) return -1;
  std::vector<Rec> d(k);
  for (u64 i = 0; i < k; ++i) d[i] = {u[i], v[i], a ? a[i] : 1.0};
  return int64_t(hit_leiden(st, d));
}

uint32_t or_level_n(void* h, int level) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return 0;
  return st.L[level].G.n;
}
int or_levels(void* h) { return static_cast<State*>(h)->P; }

int or_read_map(void* h, int level, int which, uint32_t* out) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return -1;
  auto& m = level_map(st.L[level], which);
  std::copy(m.begin(), m.end(), out);
  return int(m.size());
}

int or_write_map(void* h, int level, int which, const uint32_t* in) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return -1;
  auto& m = level_map(st.L[level], which);
  std::copy(in, in + m.size(), m.begin());
  return 0;
}

int or_read_set(void* h, int level, int which, uint8_t* out) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return -1;
  auto& s = level_set(st.L[level], which);
  u32 n = st.L[level].G.n;
  for (u32 i = 0; i < n; ++i) out[i] = i < s.size() ? s[i] : 0;
  return int(n);
}

uint32_t or_level_iters(void* h, int level) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return 0;
  return st.L[level].iters;
}

uint64_t or_read_graph(void* h, int level, uint64_t* off, uint32_t* cols, double* w) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return 0;
  const Graph& g = st.L[level].G;
  u64 nnz = 0;
  for (u32 v = 0; v < g.n; ++v) {
    if (off) off[v] = nnz;
    for (size_t j = 0; j < g.col[v].size(); ++j) {
      if (cols) cols[nnz] = g.col[v][j];
      if (w) w[nnz] = g.wt[v][j];
      ++nnz;
    }
  }
  if (off) off[g.n] = nnz;
  return nnz;
}

uint64_t or_read_delta(void* h, int level, uint32_t* u, uint32_t* v, double* w) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return 0;
  const auto& d = st.L[level].dG;
  if (u)
    for (size_t i = 0; i < d.size(); ++i) {
      u[i] = d[i].u;
      v[i] = d[i].v;
      w[i] = d[i].w;
    }
  return d.size();
}

double or_twice_m(void* h, int level) {
  State& st = *static_cast<State*>(h);
  if (level < 0 || level >= st.P) return 0;
  return st.L[level].G.twice_m;
}

int or_hierarchy(void* h, uint32_t* out) {
  State& st = *static_cast<State*>(h);
  const u32 n = st.n0;
  std::vector<u32> cur(n);
  std::iota(cur.begin(), cur.end(), 0u);
  for (int p = 0; p < st.P - 1; ++p) {
    for (u32 v = 0; v < n; ++v) cur[v] = st.L[p].s_cur[cur[v]];
    std::copy(cur.begin(), cur.end(), out + u64(p) * n);
  }
  for (u32 v = 0; v < n; ++v) out[u64(st.P - 1) * n + v] = st.L[st.P - 1].f[cur[v]];
  return 0;
}

double or_modularity_labels(void* h, const uint32_t* labels) {
  State& st = *static_cast<State*>(h);
  return modularity(st.L[0].G, labels);
}

int or_apply_delta_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew, uint64_t k,
                      const uint32_t* du, const uint32_t* dv, const double* dw, uint32_t n_new, uint64_t* out_m,
                      uint32_t* ou, uint32_t* ov, double* ow) {
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  std::vector<Rec> d(k);
  for (u64 i = 0; i < k; ++i) d[i] = {du[i], dv[i], dw ? dw[i] : 1.0};
  apply_delta(g, d, n_new);
  u64 c = 0;
  for (u32 v = 0; v < g.n; ++v)
    for (size_t j = 0; j < g.col[v].size(); ++j)
      if (g.col[v][j] >= v) {
        if (ou) {
          ou[c] = v;
          ov[c] = g.col[v][j];
          ow[c] = g.wt[v][j];
        }
        ++c;
      }
  *out_m = c;
  return int(g.n);
}

int or_inc_movement_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew, uint32_t* f,
                       const uint32_t* s, uint64_t k, const uint32_t* du, const uint32_t* dv, const double* dw,
                       double gamma, int mode, uint8_t* outB, uint8_t* outK) {
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  std::vector<u32> fv(f, f + n), sv(s, s + n);
  std::vector<Rec> d(k);
  for (u64 i = 0; i < k; ++i) d[i] = {du[i], dv[i], dw ? dw[i] : 1.0};
  MoveResult r = inc_movement(g, d, fv, sv, gamma, mode, g.n);
  std::copy(fv.begin(), fv.end(), f);
  if (outB) std::copy(r.B.begin(), r.B.end(), outB);
  if (outK) std::copy(r.K.begin(), r.K.end(), outK);
  return int(r.iters);
}

int or_inc_refinement_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                         const uint32_t* f, uint32_t* s, const uint8_t* K, double gamma, int mode, uint8_t* outR) {
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  std::vector<u32> fv(f, f + n), sv(s, s + n);
  std::vector<uint8_t> Kv(K, K + n);
  std::vector<uint8_t> R = inc_refinement(g, fv, sv, Kv, gamma, mode);
  std::copy(sv.begin(), sv.end(), s);
  if (outR) std::copy(R.begin(), R.end(), outR);
  return 0;
}

int64_t or_inc_aggregation_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                              uint64_t k, const uint32_t* du, const uint32_t* dv, const double* dw, uint32_t* s_pre,
                              const uint32_t* s_cur, const uint8_t* R, uint32_t* ou, uint32_t* ov, double* ow,
                              uint64_t cap, uint32_t* next_n) {
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  std::vector<u32> sp(s_pre, s_pre + n), sc(s_cur, s_cur + n);
  std::vector<uint8_t> Rv(R, R + n);
  std::vector<Rec> d(k);
  for (u64 i = 0; i < k; ++i) d[i] = {du[i], dv[i], dw ? dw[i] : 1.0};
  u32 nn = 0;
  std::vector<Rec> out = inc_aggregation(g, d, sp, sc, Rv, nn);
  std::copy(sp.begin(), sp.end(), s_pre);
  if (next_n) *next_n = nn;
  if (out.size() > cap) return -int64_t(out.size());
  for (size_t i = 0; i < out.size(); ++i) {
    ou[i] = out[i].u;
    ov[i] = out[i].v;
    ow[i] = out[i].w;
  }
  return int64_t(out.size());
}

int or_def_update_fn(int P, const uint32_t* sizes, uint32_t* f_flat, const uint32_t* s_flat, uint8_t* B_flat) {
  std::vector<std::vector<u32>> f(P), s(P);
  std::vector<std::vector<uint8_t>> B(P);
  u64 off = 0;
  for (int p = 0; p < P; ++p) {
    f[p].assign(f_flat + off, f_flat + off + sizes[p]);
    s[p].assign(s_flat + off, s_flat + off + sizes[p]);
    B[p].assign(B_flat + off, B_flat + off + sizes[p]);
    off += sizes[p];
  }
  std::vector<std::vector<u32>*> fm(P);
  std::vector<const std::vector<u32>*> sm(P);
  for (int p = 0; p < P; ++p) {
    fm[p] = &f[p];
    sm[p] = &s[p];
  }
  def_update(fm, sm, B);
  off = 0;
  for (int p = 0; p < P; ++p) {
    std::copy(f[p].begin(), f[p].end(), f_flat + off);
    std::copy(B[p].begin(), B[p].end(), B_flat + off);
    off += sizes[p];
  }
  return 0;
}

double or_modularity_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                        const uint32_t* labels) {
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  return modularity(g, labels);
}

uint64_t or_select_decoupled_fn(uint64_t k, const uint32_t* node, const uint32_t* from, const uint32_t* to,
                                const double* gain, uint32_t* out_nodes) {
  std::vector<MoveCand> c(k);
  for (u64 i = 0; i < k; ++i) c[i] = {node[i], from[i], to[i], 1.0, gain[i]};
  std::vector<MoveCand> sel = select_decoupled(c);
  for (size_t i = 0; i < sel.size(); ++i) out_nodes[i] = sel[i].node;
  return sel.size();
}

void or_canonicalize_fn(uint32_t n, uint32_t* labels) {
  // canonicalize_community_ids_in_place (hit_leiden.rs:358-370): first-seen contiguous ids
  std::vector<std::pair<u32, u32>> seen;  // label -> new id, kept sorted
  u32 next = 0;
  for (u32 i = 0; i < n; ++i) {
    auto it = std::lower_bound(seen.begin(), seen.end(), std::make_pair(labels[i], 0u),
                               [](const std::pair<u32, u32>& a, const std::pair<u32, u32>& b) { return a.first < b.first; });
    if (it != seen.end() && it->first == labels[i]) {
      labels[i] = it->second;
    } else {
      seen.insert(it, {labels[i], next});
      labels[i] = next++;
    }
  }
}

// ---- rand 0.8.5 StdRng + SliceRandom::shuffle (the reference loaders'
// seeded shuffle, dynamic_graph.rs:85-87; rand_core 0.6.4 seed_from_u64,
// rand_chacha 0.3.1 ChaCha12, UniformInt<u32>::sample_single_inclusive) ----
namespace {
struct ChaChaStream {
  uint32_t in[16];
  uint32_t out[16];
  int rounds;
  int pos = 16;
  static uint32_t rl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
  void next_block() {
    static const int idx[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                  {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};
    uint32_t x[16];
    for (int i = 0; i < 16; ++i) x[i] = in[i];
    for (int r = 0; r < rounds / 2; ++r)
      for (const auto& q : idx) {
        uint32_t &a = x[q[0]], &b = x[q[1]], &c = x[q[2]], &d = x[q[3]];
        a += b, d = rl(d ^ a, 16);
        c += d, b = rl(b ^ c, 12);
        a += b, d = rl(d ^ a, 8);
        c += d, b = rl(b ^ c, 7);
      }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
    uint64_t ctr = (uint64_t(in[13]) << 32 | in[12]) + 1;
    in[12] = uint32_t(ctr), in[13] = uint32_t(ctr >> 32);
    pos = 0;
  }
  uint32_t word() {
    if (pos == 16) next_block();
    return out[pos++];
  }
};
ChaChaStream chacha(const uint32_t* key, int rounds, uint64_t ctr, uint64_t stream) {
  ChaChaStream c;
  const uint32_t sigma[4] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
  for (int i = 0; i < 4; ++i) c.in[i] = sigma[i];
  for (int i = 0; i < 8; ++i) c.in[4 + i] = key[i];
  c.in[12] = uint32_t(ctr), c.in[13] = uint32_t(ctr >> 32), c.in[14] = uint32_t(stream), c.in[15] = uint32_t(stream >> 32);
  c.rounds = rounds;
  return c;
}
ChaChaStream std_rng(uint64_t seed) {
  uint32_t key[8];
  for (auto& k : key) {  // rand_core's PCG32 seed expansion
    seed = seed * 6364136223846793005ull + 11634580027462260723ull;
    const uint32_t xorshifted = uint32_t(((seed >> 18) ^ seed) >> 27), rot = uint32_t(seed >> 59);
    k = rot ? (xorshifted >> rot) | (xorshifted << (32 - rot)) : xorshifted;
  }
  return chacha(key, 12, 0, 0);
}
// gen_range(0..ubound) for u32 (rand 0.8.5): widening multiply, reject above the zone
uint32_t gen_index(ChaChaStream& r, uint32_t ubound) {
  const int lz = __builtin_clz(ubound);
  const uint32_t zone = uint32_t((uint64_t(ubound) << lz) - 1);
  for (;;) {
    const uint64_t prod = uint64_t(r.word()) * uint64_t(ubound);
    if (uint32_t(prod & 0xffffffffu) <= zone) return uint32_t(prod >> 32);
  }
}
}  // namespace

void or_stdrng_shuffle(uint64_t seed, uint64_t n, uint64_t* a) {
  for (uint64_t i = 0; i < n; ++i) a[i] = i;
  ChaChaStream r = std_rng(seed);
  for (uint64_t i = n; i >= 2; --i) std::swap(a[i - 1], a[gen_index(r, uint32_t(i))]);
}
void or_chacha_block(const uint32_t* key, int rounds, uint64_t counter, uint64_t stream, uint32_t* out) {
  ChaChaStream c = chacha(key, rounds, counter, stream);
  for (int i = 0; i < 16; ++i) out[i] = c.word();
}

// ---------------------------------------------------------------------------
// The static path: run() / multilevel_leiden (hit_leiden.rs:30-118, 167-323),
// single_level_movement (:764-932), refine_singleton_merge (:487-710),
// subcommunity_delta_q_to_empty (:712-749), aggregate_graph (:751-790).
// Canonical orders as above (sorted rows, candidates in ascending id, the
// coarse graph's edges sorted by key, the coarse initial communities numbered
// first-seen over ascending coarse id); the unseeded thread_rng of the queue
// shuffle (:856-860) is one StdRng::seed_from_u64(seed) for the whole run.
// ---------------------------------------------------------------------------
namespace {

// queue-based deterministic movement (the path the reference always takes:
// USE_PARALLEL_MULTILEVEL_MOVEMENT = false, :806)
u32 static_movement(const Graph& g, const std::vector<Rec>& delta, std::vector<u32>& f, double gamma,
                    ChaChaStream& rng) {
  const u32 n = g.n;
  std::vector<uint8_t> active(n, 0);
  for (const Rec& e : delta)
    if (e.w > 0.0 && e.u < n && e.v < n && f[e.u] != f[e.v]) active[e.u] = active[e.v] = 1;
  if (delta.empty()) std::fill(active.begin(), active.end(), 1);
  const double t = g.twice_m;
  std::vector<double> cdeg(n, 0.0);
  for (u32 i = 0; i < n; ++i) cdeg[f[i]] += g.k[i];
  std::vector<u64> order;
  for (u32 i = 0; i < n; ++i)
    if (active[i]) order.push_back(i);
  for (u64 i = order.size(); i >= 2; --i) std::swap(order[i - 1], order[gen_index(rng, u32(i))]);
  std::vector<u32> queue(order.begin(), order.end());
  std::vector<uint8_t> inq = active;
  std::vector<std::pair<u32, double>> buf;
  for (size_t h = 0; h < queue.size(); ++h) {
    const u32 v = queue[h];
    inq[v] = 0;
    const u32 cur = f[v];
    const double kv = g.k[v];
    buf.clear();
    for (size_t j = 0; j < g.col[v].size(); ++j)
      if (g.col[v][j] != v) buf.push_back({f[g.col[v][j]], g.wt[v][j]});
    std::stable_sort(buf.begin(), buf.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    double wcur = 0.0;
    for (auto& e : buf)
      if (e.first == cur) wcur += e.second;
    const double sc = cdeg[cur];
    u32 best = cur;
    double bg = 0.0;
    for (size_t i = 0; i < buf.size();) {
      const u32 c = buf[i].first;
      double wc = 0.0;
      while (i < buf.size() && buf[i].first == c) wc += buf[i++].second;
      if (c == cur) continue;
      const double gain = move_gain(wc, wcur, t, gamma, kv, sc, cdeg[c]);
      if (gain > bg) bg = gain, best = c;
    }
    if (bg <= 0.0) continue;
    f[v] = best;
    cdeg[cur] -= kv;
    cdeg[best] += kv;
    for (u32 u : g.col[v]) {
      if (u == v || f[u] == best || inq[u]) continue;
      queue.push_back(u);
      inq[u] = 1;
    }
  }
  return 1;
}

// Leiden refinement of the static path: subcommunities start as the
// communities, disconnected ones split (the largest component keeps the id,
// ties to the smaller minimum vertex), then the split-off singletons merge
// with the T-filter and the EPS tie rules of :617-693.
std::vector<u32> static_refine(const Graph& g, const std::vector<u32>& f, double gamma) {
  const u32 n = g.n;
  std::vector<u32> sub(f);
  if (n == 0) return sub;
  const double t = g.twice_m;
  if (t == 0.0) return sub;
  const double EPS = 1e-12;
  u64 maxc = 0;
  for (u32 x : f) maxc = std::max<u64>(maxc, x);
  std::vector<double> dC(maxc + 1, 0.0);
  for (u32 i = 0; i < n; ++i) dC[f[i]] += g.k[i];
  u64 next_id = maxc + 1;
  std::vector<uint8_t> refined(n, 0);
  SubIndex ix = build_members(sub, n, maxc + 1);
  std::vector<uint8_t> visited(n, 0);
  std::vector<u32> q;
  for (u64 sc = 0; sc <= maxc; ++sc) {
    const u64 b = ix.off[sc], e = ix.off[sc + 1];
    if (e - b <= 1) continue;
    std::vector<std::vector<u32>> comps;
    for (u64 pp = b; pp < e; ++pp) {
      const u32 st = ix.mem[pp];
      if (visited[st]) continue;
      q.assign(1, st);
      visited[st] = 1;
      for (size_t h = 0; h < q.size(); ++h)
        for (u32 y : g.col[q[h]])
          if (sub[y] == u32(sc) && !visited[y]) visited[y] = 1, q.push_back(y);
      comps.push_back(q);
    }
    for (u64 pp = b; pp < e; ++pp) visited[ix.mem[pp]] = 0;
    if (comps.size() <= 1) continue;
    // components are discovered from ascending start vertices, so a later
    // component of equal size has a larger minimum: the first largest is kept
    size_t keep = 0;
    for (size_t c = 1; c < comps.size(); ++c)
      if (comps[c].size() > comps[keep].size()) keep = c;
    for (size_t c = 0; c < comps.size(); ++c) {
      if (c == keep) continue;
      const u32 nid = u32(next_id++);
      for (u32 v : comps[c]) sub[v] = nid, refined[v] = 1;
    }
  }
  std::vector<u64> size(next_id, 0);
  std::vector<double> dS(next_id, 0.0);
  for (u32 v = 0; v < n; ++v) size[sub[v]] += 1, dS[sub[v]] += g.k[v];
  std::vector<u32> R;
  for (u32 v = 0; v < n; ++v)
    if (refined[v]) R.push_back(v);
  std::stable_sort(R.begin(), R.end(), [&](u32 a, u32 b) { return g.k[a] < g.k[b]; });
  std::vector<std::pair<u32, double>> buf;
  for (u32 v : R) {
    const u32 cur = sub[v];
    if (size[cur] != 1) continue;
    const u32 C = f[v];
    const double kv = g.k[v];
    buf.clear();
    double wcur = 0.0;
    for (size_t j = 0; j < g.col[v].size(); ++j) {
      const u32 u = g.col[v][j];
      if (u == v || f[u] != C) continue;
      buf.push_back({sub[u], g.wt[v][j]});
      if (sub[u] == cur) wcur += g.wt[v][j];
    }
    if (buf.empty()) continue;
    std::stable_sort(buf.begin(), buf.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    u32 best = cur;
    double bg = 0.0, bw = -1.0;
    for (size_t i = 0; i < buf.size();) {
      const u32 S = buf[i].first;
      double w = 0.0;
      while (i < buf.size() && buf[i].first == S) w += buf[i++].second;
      if (S == cur) continue;
      // subcommunity_delta_q_to_empty: d_S = 0 gives 0 (kept); w(S, C) sums
      // every row entry of every member that lies in C, a self-loop twice
      double dq = 0.0;
      if (dS[S] != 0.0) {
        double ws = 0.0;
        for (u32 i2 = 0; i2 < n; ++i2) {
          if (sub[i2] != S) continue;
          for (size_t j = 0; j < g.col[i2].size(); ++j)
            if (f[g.col[i2][j]] == C) ws += (g.col[i2][j] == i2) ? 2.0 * g.wt[i2][j] : g.wt[i2][j];
        }
        dq = (0.0 - ws) / t + gamma * dS[S] * (dC[C] - dS[S]) / (t * t);
      }
      if (!(dq <= EPS)) continue;
      const double gain = (w - wcur) / t + gamma * kv * (dS[cur] - kv - dS[S]) / (t * t);
      const bool better = gain > bg + EPS ||
                          (std::fabs(gain - bg) <= EPS && (w > bw + EPS || (std::fabs(w - bw) <= EPS && S < best)));
      if (better) best = S, bg = gain, bw = w;
    }
    if (bg > EPS) {
      sub[v] = best;
      size[cur] -= 1, size[best] += 1;
      dS[cur] -= kv, dS[best] += kv;
    }
  }
  return sub;
}

std::vector<u32> first_seen(const std::vector<u32>& a) {
  std::vector<u32> out(a.size());
  std::vector<std::pair<u32, u32>> seen;  // (label, id), sorted by label
  u32 next = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    auto it = std::lower_bound(seen.begin(), seen.end(), std::make_pair(a[i], 0u),
                               [](const auto& x, const auto& y) { return x.first < y.first; });
    if (it != seen.end() && it->first == a[i]) {
      out[i] = it->second;
    } else {
      seen.insert(it, {a[i], next});
      out[i] = next++;
    }
  }
  return out;
}

u64 count_distinct(const std::vector<u32>& a) {
  std::vector<u32> b(a);
  std::sort(b.begin(), b.end());
  return u64(std::unique(b.begin(), b.end()) - b.begin());
}

}  // namespace

// run(graph, config) (hit_leiden.rs:30-118): labels (first-seen contiguous),
// hierarchy rows (first-seen per row, at most max_levels + 1 rows, *rows set),
// returns the movement call count or -1 for bad input.
int64_t or_static_run(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew, double gamma,
                      uint32_t max_levels, uint64_t seed, uint32_t* labels, uint32_t* hier, uint32_t hier_rows,
                      uint32_t* rows) {
  if (!valid_edges(n, m, eu, ev, ew)) return -1;
  std::vector<Rec> edges(m);
  for (u64 i = 0; i < m; ++i) edges[i] = {eu[i], ev[i], ew ? ew[i] : 1.0};
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  ChaChaStream rng = std_rng(seed);
  std::vector<u32> comm(n);
  for (u32 i = 0; i < n; ++i) comm[i] = i;
  int64_t iters = static_movement(g, edges, comm, gamma, rng);
  std::vector<u32> sub = static_refine(g, comm, gamma);
  std::vector<std::vector<u32>> H;
  H.push_back(first_seen(comm));
  u64 prev = count_distinct(comm);
  for (u32 level = 1; level < max_levels; ++level) {
    if (count_distinct(sub) <= 1) break;
    // aggregate_graph: the input edges summed per (remapped) pair, an
    // intra-subcommunity edge onto a self-loop, in input order
    const std::vector<u32> remap = first_seen(sub);
    u64 cn = 0;
    for (u32 x : remap) cn = std::max<u64>(cn, u64(x) + 1);
    std::vector<std::pair<u64, double>> keyed;
    keyed.reserve(m);
    for (const Rec& e : edges) {
      const u64 a = remap[e.u], b = remap[e.v];
      keyed.push_back({(std::min(a, b) << 32) | std::max(a, b), e.w});
    }
    std::stable_sort(keyed.begin(), keyed.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    std::vector<u32> cu, cv;
    std::vector<double> cw;
    for (size_t i = 0; i < keyed.size();) {
      const u64 key = keyed[i].first;
      double w = 0.0;
      while (i < keyed.size() && keyed[i].first == key) w += keyed[i++].second;
      cu.push_back(u32(key >> 32));
      cv.push_back(u32(key & 0xffffffffu));
      cw.push_back(w);
    }
    if (cn <= 1 || cu.empty()) break;
    // each supervertex starts in its first member's community, those numbered
    // first-seen over ascending supervertex id (build_coarse_initial_partition)
    std::vector<u32> rep(cn, UINT32_MAX);
    for (u32 i = 0; i < n; ++i)
      if (rep[remap[i]] == UINT32_MAX) rep[remap[i]] = comm[i];
    std::vector<u32> cf = first_seen(rep);
    Graph cg = graph_from_edges(u32(cn), cu.size(), cu.data(), cv.data(), cw.data());
    std::vector<Rec> cedges(cu.size());
    for (size_t i = 0; i < cu.size(); ++i) cedges[i] = {cu[i], cv[i], cw[i]};
    iters += static_movement(cg, cedges, cf, gamma, rng);
    for (u32 i = 0; i < n; ++i) comm[i] = cf[remap[i]];
    H.push_back(first_seen(comm));
    const u64 now = count_distinct(comm);
    sub = static_refine(g, comm, gamma);
    if (now >= prev) break;
    prev = now;
  }
  // final polish on the input graph (:290-305)
  iters += static_movement(g, edges, comm, gamma, rng);
  std::vector<u32> fin = first_seen(comm);
  if (H.back() != fin) H.push_back(fin);
  std::copy(fin.begin(), fin.end(), labels);
  const u32 nr = std::min<u32>(u32(H.size()), hier_rows);
  for (u32 r = 0; r < nr; ++r) std::copy(H[r].begin(), H[r].end(), hier + u64(r) * n);
  *rows = u32(H.size());
  return iters;
}

// single_level_movement on its own (a fresh StdRng(seed) for the queue shuffle)
int or_static_movement_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                          uint32_t* f, double gamma, uint64_t seed) {
  if (!valid_edges(n, m, eu, ev, ew)) return -1;
  std::vector<Rec> edges(m);
  for (u64 i = 0; i < m; ++i) edges[i] = {eu[i], ev[i], ew ? ew[i] : 1.0};
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  std::vector<u32> fv(f, f + n);
  ChaChaStream rng = std_rng(seed);
  static_movement(g, edges, fv, gamma, rng);
  std::copy(fv.begin(), fv.end(), f);
  return 0;
}
// refine_singleton_merge on its own
int or_static_refine_fn(uint32_t n, uint64_t m, const uint32_t* eu, const uint32_t* ev, const double* ew,
                        const uint32_t* f, double gamma, uint32_t* sub) {
  if (!valid_edges(n, m, eu, ev, ew)) return -1;
  Graph g = graph_from_edges(n, m, eu, ev, ew);
  std::vector<u32> s2 = static_refine(g, std::vector<u32>(f, f + n), gamma);
  std::copy(s2.begin(), s2.end(), sub);
  return 0;
}

uint64_t or_empty_collisions(int reset) {
  const u64 c = g_empty_collisions;
  if (reset) g_empty_collisions = 0;
  return c;
}

int or_should_skip_aggregation_fn(uint64_t delta_len, uint32_t n, const uint8_t* R) {
  bool any = false;
  for (u32 i = 0; i < n; ++i) any = any || R[i];
  return delta_len == 0 && !any;
}

}  // extern "C"
