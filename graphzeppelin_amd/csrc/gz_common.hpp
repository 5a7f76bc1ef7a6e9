// gz_common.hpp -- types, geometry and the XXH3 short-input hash shared by host and device code.
//
// Geometry follows include/sketch.h:62-76,181,187-190 and src/sketch.cpp:8-21 of the reference;
// the hash is XXH3_64bits_withSeed on the 8 little-endian bytes of a 64-bit index
// (include/types.h:8-9, include/bucket.h:54-63), restated from the public XXH3 specification.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace gz {

// ---- XXH3 (len == 8) ---------------------------------------------------------------------------
// For an 8-byte input the seed only enters through one 64-bit constant, the "bitflip":
//   seed' = seed ^ (bswap32(lo32(seed)) << 32);  bitflip = (secret[8:16] ^ secret[16:24]) - seed'
// so each column's bitflip is computed once on the host and the per-index work is
//   h = rot32(idx) ^ bitflip; rrmxmx(h, 8).
constexpr uint64_t kPrimeMx2 = 0x9FB21C651E98DF25ULL;
constexpr uint64_t kSecretMix = 0x1cad21f72c81017cULL ^ 0xdb979083e96dd4deULL;

__host__ __device__ inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

__host__ __device__ inline uint64_t xxh3_bitflip(uint64_t seed) {
  uint32_t lo = (uint32_t)seed;
  uint32_t sw = (lo >> 24) | ((lo >> 8) & 0xff00u) | ((lo << 8) & 0xff0000u) | (lo << 24);
  seed ^= (uint64_t)sw << 32;
  return kSecretMix - seed;
}

// rot32(idx): the input is read as two little-endian 32-bit words and recombined hi + (lo << 32)
__host__ __device__ inline uint64_t xxh3_prep(uint64_t idx) { return (idx >> 32) | (idx << 32); }

__host__ __device__ inline uint64_t xxh3_mix(uint64_t prepped, uint64_t bitflip) {
  uint64_t h = prepped ^ bitflip;
  h ^= rotl64(h, 49) ^ rotl64(h, 24);
  h *= kPrimeMx2;
  h ^= (h >> 35) + 8;
  h *= kPrimeMx2;
  return h ^ (h >> 28);
}

__host__ __device__ inline uint64_t xxh3_len8(uint64_t idx, uint64_t seed) {
  return xxh3_mix(xxh3_prep(idx), xxh3_bitflip(seed));
}

// ---- pairing (src/util.cpp:45-57) -------------------------------------------------------------
__host__ __device__ inline uint64_t concat_pairing(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  return ((uint64_t)lo << 32) | hi;
}

// ---- geometry (host only: double arithmetic exactly as the reference does it) -----------------
inline uint64_t calc_vector_length(uint32_t n) { return (uint64_t)std::ceil(double(n) * (n - 1) / 2); }
inline size_t calc_cc_samples(uint32_t n, double f) {
  const double div = 1 - std::log2(2 - 0.8);
  return std::max(size_t(18), (size_t)std::ceil(f * std::log2(n) / div));
}
inline size_t calc_bkt_per_col(uint64_t vector_len) { return (size_t)(std::ceil(std::log2((double)vector_len)) + 1); }

// Geometry of one sketch family, passed by value to kernels.
struct Geo {
  uint32_t n;        // vertices (or 1 for a standalone sketch)
  uint32_t D;        // buckets per column (bkt_per_col)
  uint32_t C;        // columns (num_samples * cols_per_sample)
  uint32_t cps;      // cols_per_sample
  uint32_t samples;  // num_samples
  uint32_t nb;       // buckets per sketch = C*D + 1 (deterministic bucket last)
  uint32_t row;      // bytes between consecutive sketches in the slab (RowLayout::row)
  uint32_t low;      // depths of a column kept in the paired low region (RowLayout::low)
  uint32_t ovf;      // byte offset of the overflow region in a row (RowLayout::ovf)
  uint32_t opitch;   // bytes between consecutive columns in the overflow region (RowLayout::opitch)
  uint32_t Dp;       // LDS column pitch (D rounded up to 1 mod 16 to spread LDS banks)
  uint64_t seed;     // sketch seed (checksum seed)
  uint64_t chk_bf;   // bitflip of the checksum seed
};

inline uint32_t lds_pitch(uint32_t D) { return D + ((17u - (D % 16u)) % 16u); }

// Row of one vertex sketch in the slab.  An update touches one record in every column, at depth d
// with probability 2^-(d+1), and the deterministic bucket, so the low depths are what every flush
// reads and writes.  The low region pairs columns on 128-B lines: depths d < low of column c sit at
// byte (c / 2) * 128 + (c % 2) * 64 + d * rb (low = the records that fit in 64 B: 5 of 12 B, 4 of
// 16 B), and the deterministic bucket takes slot C of the same pairing (the free half of the last
// line when C is odd).  Depths d >= low of column c follow in the overflow region at byte
// ovf + c * opitch + (d - low) * rb, each column's run starting on a 128-B line, so a Borůvka round
// reads column r as one low half-line and ceil((D - low) * rb / 128) overflow lines.
struct RowLayout {
  uint32_t low, ovf, opitch, row;
};
inline RowLayout row_layout(uint32_t C, uint32_t D, uint32_t rb) {
  RowLayout L{};
  L.low = std::min(D, 64u / rb);
  L.ovf = (C + 2) / 2 * 128u;  // ceil((C + 1) / 2) lines: C columns and the deterministic bucket
  L.opitch = D > L.low ? ((D - L.low) * rb + 127u) & ~127u : 0u;
  L.row = (L.ovf + C * L.opitch + 255u) & ~255u;
  return L;
}

// Per-tile work descriptor of the update kernel: `count` destinations of `vertex` starting at
// `begin` in the destination array; `split` = the vertex has other tiles in the same launch, so
// the write-back must use atomics.
struct Tile {
  uint32_t vertex, begin, count, split;
};

enum SampleCode : int { kGood = 0, kZero = 1, kFail = 2 };

}  // namespace gz
