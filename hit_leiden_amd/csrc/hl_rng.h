// hl_rng.h -- the seeded shuffle of the reference's loaders and static path:
// rand 0.8.5 `StdRng::seed_from_u64(seed)` followed by `SliceRandom::shuffle`
// (src/benchmark/dynamic_graph.rs:47-49,85-87; the static movement queue,
// src/core/algorithm/hit_leiden.rs:856-860, shuffles with an unseeded
// thread_rng, which this replaces by the same seeded generator).
//
// The published algorithms, restated (rand 0.8.5, rand_core 0.6.4,
// rand_chacha 0.3.1, all pinned by the reference's Cargo.lock):
//  * seed_from_u64: a PCG32 stream (multiplier 6364136223846793005, increment
//    11634580027462260723, state advanced before each output, output
//    rotr(((s >> 18) ^ s) >> 27, s >> 59)) fills the 32-byte seed with
//    little-endian words;
//  * StdRng = ChaCha with 12 rounds, key = seed, 64-bit block counter from 0,
//    64-bit stream id 0; next_u32 returns the keystream words in order (four
//    blocks per refill, block after block);
//  * gen_range(0..n) for u32: range r = n, zone = (r << clz(r)) - 1, draw
//    v = next_u32 until the low half of v * r is <= zone, return its high half;
//  * shuffle: for i = len-1 down to 1, swap(i, gen_index(i + 1)), where
//    gen_index takes the u32 path for bounds up to 2^32 - 1.
#pragma once

#include <cstdint>
#include <cstring>

namespace hl {

class StdRng {
 public:
  explicit StdRng(uint64_t seed) {
    uint32_t key[8];
    uint64_t st = seed;
    for (int i = 0; i < 8; ++i) key[i] = pcg32(st);
    init(key, 12);
  }
  // ChaCha with `rounds` rounds over a 256-bit key (little-endian words),
  // block counter `counter` and 64-bit stream id `stream` (test vectors)
  StdRng(const uint32_t key[8], int rounds, uint64_t counter = 0, uint64_t stream = 0) {
    init(key, rounds);
    s_[12] = uint32_t(counter);
    s_[13] = uint32_t(counter >> 32);
    s_[14] = uint32_t(stream);
    s_[15] = uint32_t(stream >> 32);
  }
  uint32_t next_u32() {
    if (idx_ >= 64) refill();
    return buf_[idx_++];
  }
  // rand 0.8.5 UniformInt<u32>::sample_single_inclusive(0, n - 1), n >= 1
  uint32_t below(uint32_t n) {
    const uint32_t zone = (n << __builtin_clz(n)) - 1u;
    while (true) {
      const uint64_t m = uint64_t(next_u32()) * n;
      if (uint32_t(m) <= zone) return uint32_t(m >> 32);
    }
  }
  // SliceRandom::shuffle (indices up to 2^32 - 1)
  template <class T>
  void shuffle(T* a, uint64_t len) {
    for (uint64_t i = len; i-- > 1;) {
      const uint64_t j = below(uint32_t(i + 1));
      T t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  // one ChaCha block at the current counter into out[16] (test vectors); advances the counter
  void block(uint32_t out[16]) {
    uint32_t x[16];
    std::memcpy(x, s_, sizeof(x));
    for (int r = 0; r < rounds_; r += 2) {
      qr(x[0], x[4], x[8], x[12]);
      qr(x[1], x[5], x[9], x[13]);
      qr(x[2], x[6], x[10], x[14]);
      qr(x[3], x[7], x[11], x[15]);
      qr(x[0], x[5], x[10], x[15]);
      qr(x[1], x[6], x[11], x[12]);
      qr(x[2], x[7], x[8], x[13]);
      qr(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s_[i];
    if (++s_[12] == 0) ++s_[13];
  }

 private:
  uint32_t s_[16];
  uint32_t buf_[64];
  int idx_ = 64;
  int rounds_ = 12;

  static uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
  static void qr(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    a += b; d ^= a; d = rotl(d, 16);
    c += d; b ^= c; b = rotl(b, 12);
    a += b; d ^= a; d = rotl(d, 8);
    c += d; b ^= c; b = rotl(b, 7);
  }
  static uint32_t pcg32(uint64_t& st) {
    st = st * 6364136223846793005ull + 11634580027462260723ull;
    const uint32_t xs = uint32_t(((st >> 18) ^ st) >> 27);
    const uint32_t rot = uint32_t(st >> 59);
    return (xs >> rot) | (xs << ((32 - rot) & 31));
  }
  void init(const uint32_t key[8], int rounds) {
    rounds_ = rounds;
    s_[0] = 0x61707865u;  // "expand 32-byte k"
    s_[1] = 0x3320646eu;
    s_[2] = 0x79622d32u;
    s_[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i) s_[4 + i] = key[i];
    s_[12] = s_[13] = s_[14] = s_[15] = 0;
    idx_ = 64;
  }
  void refill() {
    for (int b = 0; b < 4; ++b) block(buf_ + 16 * b);
    idx_ = 0;
  }
};

}  // namespace hl
