// hl_gen.cpp -- seeded synthetic graphs and edge-delta streams for the five
// workloads (LFR-style, planted partition, R-MAT, SBM) and their batches.
// PRNG: splitmix64-seeded xoshiro256**.  Plain C++ (host only).  Also the
// reference loaders' seeded shuffle (hl_rng.h: StdRng + SliceRandom::shuffle).
#include "hl_rng.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace {

using u32 = uint32_t;
using u64 = uint64_t;

struct Rng {
  u64 s[4];
  static u64 splitmix(u64& x) {
    u64 z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  explicit Rng(u64 seed) {
    u64 x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }
  u64 next() {
    u64 r = rotl(s[1] * 5, 7) * 9;
    u64 t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uni() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  u64 below(u64 n) { return n ? next() % n : 0; }
  double normal() {
    double a = uni(), b = uni();
    if (a < 1e-300) a = 1e-300;
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
  }
};

inline u64 key(u32 a, u32 b) {
  if (a > b) std::swap(a, b);
  return (u64(a) << 32) | b;
}

struct EdgeSet {
  u32 n = 0;
  std::vector<u64> keys;      // sorted canonical keys
  std::vector<double> w;      // weight per key
  std::vector<u32> labels;    // planted community of each vertex (if any)
  // batches: concatenated
  std::vector<u32> bu, bv;
  std::vector<double> ba;
  std::vector<u64> boff{0};
};

void finalize(EdgeSet& es, std::vector<u64>& k, bool weighted, Rng& rng, double wlo, double whi) {
#ifdef _OPENMP
  __gnu_parallel::sort(k.begin(), k.end());
#else
  std::sort(k.begin(), k.end());
#endif
  k.erase(std::unique(k.begin(), k.end()), k.end());
  es.keys.swap(k);
  es.w.assign(es.keys.size(), 1.0);
  if (weighted)
    for (auto& x : es.w) x = wlo + (whi - wlo) * rng.uni();
}

// power-law sample in [lo, hi] with exponent tau (continuous, inverse CDF)
double powerlaw(Rng& r, double lo, double hi, double tau) {
  double a = std::pow(lo, 1.0 - tau), b = std::pow(hi, 1.0 - tau);
  double u = r.uni();
  return std::pow(a + u * (b - a), 1.0 / (1.0 - tau));
}

// LFR-style benchmark (degrees ~ k^-tau1, community sizes ~ s^-tau2 in
// [cmin,cmax], mixing mu), trimmed or topped up to exactly target_m edges.
EdgeSet* lfr(u32 n, double tau1, double mean_deg, double tau2, u32 cmin, u32 cmax, double mu, u64 target_m,
             u64 seed) {
  Rng rng(seed);
  EdgeSet* es = new EdgeSet();
  es->n = n;
  const double kmax = std::max(8.0, std::sqrt(double(n)) * 2.0);
  // find kmin so that the mean matches
  double lo = 0.5, hi = mean_deg;
  for (int it = 0; it < 60; ++it) {
    double mid = 0.5 * (lo + hi);
    // mean of continuous truncated power law
    double a = 1.0 - tau1, b = 2.0 - tau1;
    double m = (a / b) * (std::pow(kmax, b) - std::pow(mid, b)) / (std::pow(kmax, a) - std::pow(mid, a));
    if (m < mean_deg)
      lo = mid;
    else
      hi = mid;
  }
  const double kmin = 0.5 * (lo + hi);
  std::vector<u32> deg(n);
  for (u32 v = 0; v < n; ++v) deg[v] = std::max<u32>(1, u32(std::lround(powerlaw(rng, kmin, kmax, tau1))));
  // community sizes
  std::vector<u32> sizes;
  u64 tot = 0;
  while (tot < n) {
    u32 s = u32(std::lround(powerlaw(rng, cmin, cmax, tau2)));
    s = std::min<u32>(std::max<u32>(s, cmin), cmax);
    if (tot + s > n) s = u32(n - tot);
    sizes.push_back(s);
    tot += s;
  }
  if (sizes.size() > 1 && sizes.back() < cmin) {
    sizes[sizes.size() - 2] += sizes.back();
    sizes.pop_back();
  }
  // assign vertices, highest degree first, to communities with room
  std::vector<u32> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](u32 a, u32 b) { return deg[a] > deg[b]; });
  std::vector<u32> room(sizes), comm(n);
  std::vector<u32> byroom(sizes.size());
  std::iota(byroom.begin(), byroom.end(), 0u);
  for (u32 v : order) {
    u32 need = u32(std::lround((1.0 - mu) * deg[v])) + 1;
    u32 pick = u32(-1);
    for (int tries = 0; tries < 64; ++tries) {
      u32 c = u32(rng.below(sizes.size()));
      if (room[c] > 0 && sizes[c] >= need) {
        pick = c;
        break;
      }
    }
    if (pick == u32(-1))
      for (u32 c = 0; c < sizes.size(); ++c)
        if (room[c] > 0 && (pick == u32(-1) || sizes[c] > sizes[pick])) pick = c;
    comm[v] = pick;
    room[pick]--;
  }
  std::vector<std::vector<u32>> stubs_in(sizes.size());
  std::vector<u32> stubs_out;
  for (u32 v = 0; v < n; ++v) {
    u32 kin = u32(std::lround((1.0 - mu) * deg[v]));
    kin = std::min<u32>(kin, sizes[comm[v]] - 1);
    for (u32 i = 0; i < kin; ++i) stubs_in[comm[v]].push_back(v);
    for (u32 i = kin; i < deg[v]; ++i) stubs_out.push_back(v);
  }
  std::vector<u64> keys;
  auto pair_up = [&](std::vector<u32>& st) {
    for (size_t i = st.size(); i > 1; --i) std::swap(st[i - 1], st[rng.below(i)]);
    for (size_t i = 0; i + 1 < st.size(); i += 2)
      if (st[i] != st[i + 1]) keys.push_back(key(st[i], st[i + 1]));
  };
  for (auto& st : stubs_in) pair_up(st);
  pair_up(stubs_out);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  // exact edge count
  if (target_m) {
    while (keys.size() > target_m) {
      size_t i = rng.below(keys.size());
      keys[i] = keys.back();
      keys.pop_back();
    }
    std::unordered_set<u64> have(keys.begin(), keys.end());
    std::vector<std::vector<u32>> members(sizes.size());
    for (u32 v = 0; v < n; ++v) members[comm[v]].push_back(v);
    while (keys.size() < target_m) {
      u32 a = u32(rng.below(n)), b;
      if (rng.uni() < 1.0 - mu) {
        auto& m = members[comm[a]];
        b = m[rng.below(m.size())];
      } else {
        b = u32(rng.below(n));
      }
      if (a == b) continue;
      u64 k = key(a, b);
      if (have.insert(k).second) keys.push_back(k);
    }
  }
  finalize(*es, keys, false, rng, 1, 1);
  es->labels = comm;
  return es;
}

// planted partition: blocks with log-normal sizes, mean degree, mixing mu,
// weights uniform in [wlo, whi]
EdgeSet* planted(u32 n, u32 blocks, double sigma, double mean_deg, double mu, double wlo, double whi, u64 seed) {
  Rng rng(seed);
  EdgeSet* es = new EdgeSet();
  es->n = n;
  std::vector<double> raw(blocks);
  double sum = 0;
  for (auto& x : raw) {
    x = std::exp(sigma * rng.normal());
    sum += x;
  }
  std::vector<u32> sizes(blocks);
  u64 tot = 0;
  for (u32 b = 0; b < blocks; ++b) {
    sizes[b] = std::max<u32>(2, u32(raw[b] / sum * n));
    tot += sizes[b];
  }
  // fix the total
  while (tot > n) {
    u32 b = u32(rng.below(blocks));
    if (sizes[b] > 2) sizes[b]--, tot--;
  }
  while (tot < n) sizes[rng.below(blocks)]++, tot++;
  std::vector<u32> start(blocks + 1, 0);
  for (u32 b = 0; b < blocks; ++b) start[b + 1] = start[b] + sizes[b];
  std::vector<u32> lab(n);
  for (u32 b = 0; b < blocks; ++b)
    for (u32 v = start[b]; v < start[b + 1]; ++v) lab[v] = b;
  std::vector<u64> keys;
  keys.reserve(u64(n * mean_deg / 2 * 1.02));
  for (u32 b = 0; b < blocks; ++b) {
    u64 s = sizes[b];
    u64 want = std::min<u64>(u64(std::llround(s * mean_deg * (1.0 - mu) / 2.0)), s * (s - 1) / 2);
    std::unordered_set<u64> have;
    have.reserve(want * 2);
    while (have.size() < want) {
      u32 a = start[b] + u32(rng.below(s)), c = start[b] + u32(rng.below(s));
      if (a != c) have.insert(key(a, c));
    }
    keys.insert(keys.end(), have.begin(), have.end());
  }
  u64 want_out = u64(std::llround(double(n) * mean_deg * mu / 2.0));
  std::vector<u64> out;
  out.reserve(want_out * 11 / 10);
  while (true) {
    while (out.size() < want_out + want_out / 50 + 16) {
      u32 a = u32(rng.below(n)), c = u32(rng.below(n));
      if (lab[a] != lab[c]) out.push_back(key(a, c));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    if (out.size() >= want_out) break;
  }
  // keep a random subset of exactly want_out
  for (size_t i = out.size(); i > 1; --i) std::swap(out[i - 1], out[rng.below(i)]);
  out.resize(want_out);
  keys.insert(keys.end(), out.begin(), out.end());
  finalize(*es, keys, true, rng, wlo, whi);
  es->labels = lab;
  return es;
}

// R-MAT (a, b, c, d = 1-a-b-c), 2^scale vertices, ef * 2^scale samples,
// symmetrised, deduplicated, self-loops dropped, unit weights
EdgeSet* rmat(u32 scale, u32 ef, double a, double b, double c, u64 seed, int threads) {
  EdgeSet* es = new EdgeSet();
  const u64 n = 1ull << scale;
  es->n = u32(n);
  const u64 m = u64(ef) << scale;
  std::vector<u64> keys(m);
  int T = 1;
#ifdef _OPENMP
  T = threads > 0 ? threads : omp_get_max_threads();
#endif
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel num_threads(T)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    Rng rng(seed * 1000003ull + u64(t) * 7919ull + 17);
    u64 lo = m * u64(t) / u64(T), hi = m * u64(t + 1) / u64(T);
    for (u64 i = lo; i < hi; ++i) {
      u64 u = 0, v = 0;
      for (u32 l = 0; l < scale; ++l) {
        double r = rng.uni();
        u64 bu = 0, bv = 0;
        if (r < a) {
        } else if (r < ab) {
          bv = 1;
        } else if (r < abc) {
          bu = 1;
        } else {
          bu = 1;
          bv = 1;
        }
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      keys[i] = (u == v) ? ~0ull : key(u32(u), u32(v));
    }
  }
  Rng rng(seed);
  finalize(*es, keys, false, rng, 1, 1);
  if (!es->keys.empty() && es->keys.back() == ~0ull) {
    es->keys.pop_back();
    es->w.pop_back();
  }
  return es;
}

// stochastic block model with `blocks` equal-ish blocks
EdgeSet* sbm(u32 n, u32 blocks, double mean_deg, double mu, u64 seed) {
  Rng rng(seed);
  EdgeSet* es = new EdgeSet();
  es->n = n;
  std::vector<u32> lab(n);
  for (u32 v = 0; v < n; ++v) lab[v] = u32(u64(v) * blocks / n);
  u64 m = u64(std::llround(n * mean_deg / 2.0));
  std::vector<u64> keys;
  std::unordered_set<u64> have;
  have.reserve(m * 2);
  std::vector<u32> start(blocks + 1);
  for (u32 b = 0; b <= blocks; ++b) start[b] = u32((u64(b) * n + blocks - 1) / blocks);
  u64 guard = 0;
  while (have.size() < m && guard++ < 50 * m + 1000) {
    u32 a = u32(rng.below(n)), c;
    if (rng.uni() < 1.0 - mu) {
      u32 bl = lab[a];
      u32 lo = start[bl], hi = start[bl + 1];
      if (hi <= lo + 1) continue;
      c = lo + u32(rng.below(hi - lo));
    } else {
      c = u32(rng.below(n));
    }
    if (a != c) have.insert(key(a, c));
  }
  keys.assign(have.begin(), have.end());
  finalize(*es, keys, false, rng, 1, 1);
  es->labels = lab;
  return es;
}

struct Live {
  const EdgeSet& es;
  std::unordered_set<u64> added, removed;
  explicit Live(const EdgeSet& e) : es(e) {}
  bool exists(u64 k) const {
    if (added.count(k)) return true;
    if (removed.count(k)) return false;
    return std::binary_search(es.keys.begin(), es.keys.end(), k);
  }
};

// endpoint modes: 0 degree-proportional, 1 uniform, 2 R-MAT sampled
u64 rmat_vertex_pair(Rng& rng, u32 scale, double a, double b, double c, u32& u, u32& v) {
  const double ab = a + b, abc = a + b + c;
  u64 x = 0, y = 0;
  for (u32 l = 0; l < scale; ++l) {
    double r = rng.uni();
    u64 bu = 0, bv = 0;
    if (r < a) {
    } else if (r < ab) {
      bv = 1;
    } else if (r < abc) {
      bu = 1;
    } else {
      bu = bv = 1;
    }
    x = (x << 1) | bu;
    y = (y << 1) | bv;
  }
  u = u32(x);
  v = u32(y);
  return 0;
}

void make_batches(EdgeSet& es, u32 nb, u64 ins, u64 del, int mode, double wlo, double whi, u32 scale, double a,
                  double b, double c, u64 seed) {
  Rng rng(seed);
  Live live(es);
  std::unordered_map<u64, double> added_w;
  es.bu.clear();
  es.bv.clear();
  es.ba.clear();
  es.boff.assign(1, 0);
  const bool weighted = wlo != 1.0 || whi != 1.0;
  for (u32 bi = 0; bi < nb; ++bi) {
    std::unordered_set<u64> inbatch;
    u64 got = 0, guard = 0;
    while (got < ins && guard++ < 100 * ins + 1000) {
      u32 u, v;
      if (mode == 0) {
        // degree-proportional: endpoint of a random existing edge
        u64 e1 = es.keys[rng.below(es.keys.size())], e2 = es.keys[rng.below(es.keys.size())];
        u = (rng.next() & 1) ? u32(e1 >> 32) : u32(e1);
        v = (rng.next() & 1) ? u32(e2 >> 32) : u32(e2);
      } else if (mode == 1) {
        u = u32(rng.below(es.n));
        v = u32(rng.below(es.n));
      } else {
        rmat_vertex_pair(rng, scale, a, b, c, u, v);
      }
      if (u == v) continue;
      u64 k = key(u, v);
      if (live.exists(k) || inbatch.count(k)) continue;
      inbatch.insert(k);
      double w = weighted ? wlo + (whi - wlo) * rng.uni() : 1.0;
      es.bu.push_back(u);
      es.bv.push_back(v);
      es.ba.push_back(w);
      added_w[k] = w;
      ++got;
    }
    u64 dgot = 0;
    guard = 0;
    while (dgot < del && guard++ < 100 * del + 1000) {
      u64 i = rng.below(es.keys.size());
      u64 k = es.keys[i];
      if (live.removed.count(k) || inbatch.count(k)) continue;
      inbatch.insert(k);
      es.bu.push_back(u32(k >> 32));
      es.bv.push_back(u32(k));
      es.ba.push_back(-es.w[i]);
      live.removed.insert(k);
      ++dgot;
    }
    for (u64 i = es.boff.back(); i < es.bu.size(); ++i)
      if (es.ba[i] > 0) live.added.insert(key(es.bu[i], es.bv[i]));
    es.boff.push_back(es.bu.size());
  }
}

}  // namespace

extern "C" {

void* hg_lfr(uint32_t n, double tau1, double mean_deg, double tau2, uint32_t cmin, uint32_t cmax, double mu,
             uint64_t target_m, uint64_t seed) {
  return lfr(n, tau1, mean_deg, tau2, cmin, cmax, mu, target_m, seed);
}
void* hg_planted(uint32_t n, uint32_t blocks, double sigma, double mean_deg, double mu, double wlo, double whi,
                 uint64_t seed) {
  return planted(n, blocks, sigma, mean_deg, mu, wlo, whi, seed);
}
void* hg_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, int threads) {
  return rmat(scale, ef, a, b, c, seed, threads);
}
void* hg_sbm(uint32_t n, uint32_t blocks, double mean_deg, double mu, uint64_t seed) {
  return sbm(n, blocks, mean_deg, mu, seed);
}
uint32_t hg_n(void* h) { return static_cast<EdgeSet*>(h)->n; }
uint64_t hg_m(void* h) { return static_cast<EdgeSet*>(h)->keys.size(); }
void hg_edges(void* h, uint32_t* u, uint32_t* v, double* w) {
  EdgeSet& es = *static_cast<EdgeSet*>(h);
  for (size_t i = 0; i < es.keys.size(); ++i) {
    u[i] = u32(es.keys[i] >> 32);
    v[i] = u32(es.keys[i]);
    if (w) w[i] = es.w[i];
  }
}
int hg_labels(void* h, uint32_t* out) {
  EdgeSet& es = *static_cast<EdgeSet*>(h);
  if (es.labels.empty()) return 0;
  std::copy(es.labels.begin(), es.labels.end(), out);
  return 1;
}
void hg_make_batches(void* h, uint32_t nb, uint64_t ins, uint64_t del, int mode, double wlo, double whi,
                     uint32_t scale, double a, double b, double c, uint64_t seed) {
  make_batches(*static_cast<EdgeSet*>(h), nb, ins, del, mode, wlo, whi, scale, a, b, c, seed);
}
uint32_t hg_num_batches(void* h) { return uint32_t(static_cast<EdgeSet*>(h)->boff.size() - 1); }
uint64_t hg_batch_len(void* h, uint32_t i) {
  EdgeSet& es = *static_cast<EdgeSet*>(h);
  return es.boff[i + 1] - es.boff[i];
}
void hg_batch(void* h, uint32_t i, uint32_t* u, uint32_t* v, double* a) {
  EdgeSet& es = *static_cast<EdgeSet*>(h);
  u64 s = es.boff[i], e = es.boff[i + 1];
  std::copy(es.bu.begin() + s, es.bu.begin() + e, u);
  std::copy(es.bv.begin() + s, es.bv.begin() + e, v);
  std::copy(es.ba.begin() + s, es.ba.begin() + e, a);
}
void hg_free(void* h) { delete static_cast<EdgeSet*>(h); }

// StdRng::seed_from_u64(seed) + SliceRandom::shuffle of [0, n): out[i] is the
// index of the element that lands at position i (dynamic_graph.rs:85-87)
void hg_shuffle_order(uint64_t seed, uint64_t n, uint64_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = i;
  hl::StdRng r(seed);
  r.shuffle(out, n);
}
// the first count next_u32 outputs of StdRng::seed_from_u64(seed)
void hg_stdrng_u32(uint64_t seed, uint64_t count, uint32_t* out) {
  hl::StdRng r(seed);
  for (uint64_t i = 0; i < count; ++i) out[i] = r.next_u32();
}
// one ChaCha block (key words little-endian, 64-bit counter and stream id)
void hg_chacha_block(const uint32_t* key, int rounds, uint64_t counter, uint64_t stream, uint32_t* out) {
  hl::StdRng r(key, rounds, counter, stream);
  r.block(out);
}

}  // extern "C"
