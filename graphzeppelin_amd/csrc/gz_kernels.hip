// gz_kernels.hip -- hand-written gfx950 kernels for GraphZeppelin's streaming-CC hot path.
//
// Data layout in HBM (one "slab" per algorithm instance): row v holds vertex v's sketch as packed
// bucket records {u64 alpha; G gamma} (G = u32 for the 12-byte reference bucket, u64 for the
// 16-byte one).  The low depths of two columns share a 128-B line (with the deterministic bucket in
// the pairing's last slot) and deeper depths follow per column in an overflow region (RowLayout in
// gz_common.hpp).  Kernels address buckets by the reference's index b = c*D + d (deterministic
// bucket last, b = C*D, src/sketch.cpp:69-83), so packing a row reproduces the reference's image.
// Everything is integer work; nothing here is a dense contraction, so no MFMA is used.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "gz_kernels.hpp"

namespace gz {

namespace {

constexpr int kWave = 64;

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint32_t rd_lane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ inline uint64_t rd_lane_u64(uint64_t v, int lane) {
  uint32_t lo = rd_lane_u32((uint32_t)v, lane), hi = rd_lane_u32((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = __shfl((unsigned)(uint32_t)v, src), hi = __shfl((unsigned)(uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int o) {
  uint32_t lo = __shfl_xor((unsigned)(uint32_t)v, o), hi = __shfl_xor((unsigned)(uint32_t)(v >> 32), o);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t wave_xor_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((unsigned)(uint32_t)v, o), hi = __shfl_xor((unsigned)(uint32_t)(v >> 32), o);
    v ^= ((uint64_t)hi << 32) | lo;
  }
  return v;
}
// XOR over the wave of a 32-bit value, returned to every lane: DPP row shifts build each row's XOR
// in its lane 15, two row broadcasts carry it to lane 63 (no LDS round trips).  Every lane must be
// active.
__device__ inline uint32_t wave_xor_dpp_u32(uint32_t v) {
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1);
}
__device__ inline uint64_t wave_xor_dpp_u64(uint64_t v) {
  const uint32_t lo = wave_xor_dpp_u32((uint32_t)v), hi = wave_xor_dpp_u32((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t wave_incl_scan_u64(uint64_t x) {
  const uint32_t lane = lane_id();
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t lo = __shfl_up((unsigned)(uint32_t)x, o), hi = __shfl_up((unsigned)(uint32_t)(x >> 32), o);
    uint64_t y = ((uint64_t)hi << 32) | lo;
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

template <class T>
__device__ inline void lds_xor(T* p, T v) {
  __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ inline void gbl_xor(T* p, T v) {
  __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void gbl_or(T* p, T v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T ld_acq(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void st_rlx(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- the slab: packed bucket records -----------------------------------------------------------
// Every bucket is the record {u64 alpha; G gamma} -- the reference's packed Bucket, 12 or 16 bytes.
// Row v of the slab holds sketch v (layout in gz_common.hpp, RowLayout): depths below g.low of two
// columns share a 128-B line, with the deterministic bucket in the pairing's slot C; deeper depths
// of each column follow in the overflow region, each column's run on its own lines.  A 12-byte
// record is only 4-byte aligned, so it is moved as 32-bit words, and a 64-bit XOR atomic on it
// becomes two 32-bit ones (XOR acts bit by bit, so the result is the same).
// rec_col_off: byte of (column c, depth d) within a row; det_col_off: byte of the deterministic
// bucket; rec_off: byte of the reference's bucket index b (c * D + d, the deterministic bucket last)
// of sketch v.
template <class G>
__device__ inline uint32_t rec_col_off(const Geo& g, uint32_t c, uint32_t d) {
  constexpr uint32_t RB = 8 + sizeof(G);
  return d < g.low ? (c >> 1) * 128u + (c & 1u) * 64u + d * RB : g.ovf + c * g.opitch + (d - g.low) * RB;
}
__device__ inline uint32_t det_col_off(const Geo& g) { return (g.C >> 1) * 128u + (g.C & 1u) * 64u; }
// Deep flags (rows of 12-byte records only): the word after a column's at most five low records,
// byte 60 of its half-line, becomes 1 once any record of the column that ends past the first line
// of its overflow run (depth >= low + 10) has been written nonzero: every writer of such a record
// (the update kernel's deep path, k_packed) stores 1 there, and only a reset of the row clears it.
// A Borůvka reduce reads those records only for members whose flag is set; for the others they are
// zero.  An update reaches such a depth once in 2^15 (update, column) pairs, so most vertices never
// set their flags, and the reduce reads one overflow line of the column instead of three.
template <class G>
constexpr bool kDeepFlags = sizeof(G) == 4;
__device__ inline uint32_t deep_flag_off(uint32_t c) { return (c >> 1) * 128u + (c & 1u) * 64u + 60u; }
__device__ inline bool rec_is_deep(const Geo& g, uint32_t d) { return d >= g.low && (d - g.low) * 12u + 11u >= 128u; }
template <class G>
__device__ inline size_t rec_off(const Geo& g, uint32_t v, uint32_t b) {
  if (b == g.nb - 1) return (size_t)v * g.row + det_col_off(g);
  const uint32_t c = b / g.D, d = b - c * g.D;
  return (size_t)v * g.row + rec_col_off<G>(g, c, d);
}
template <class G>
__device__ inline void ld_rec(const unsigned char* p, uint64_t* a, G* gm) {
  if (sizeof(G) == 4) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
    *a = ((uint64_t)w[1] << 32) | w[0];
    *gm = (G)w[2];
  } else {
    const uint64_t* w = reinterpret_cast<const uint64_t*>(p);
    *a = w[0];
    *gm = (G)w[1];
  }
}
template <class G>
__device__ inline void st_rec(unsigned char* p, uint64_t a, G gm) {
  if (sizeof(G) == 4) {
    uint32_t* w = reinterpret_cast<uint32_t*>(p);
    w[0] = (uint32_t)a;
    w[1] = (uint32_t)(a >> 32);
    w[2] = (uint32_t)gm;
  } else {
    uint64_t* w = reinterpret_cast<uint64_t*>(p);
    w[0] = a;
    w[1] = (uint64_t)gm;
  }
}
template <class G>
__device__ inline void xor_rec_atomic(unsigned char* p, uint64_t a, G gm) {
  if (sizeof(G) == 4) {
    uint32_t* w = reinterpret_cast<uint32_t*>(p);
    if ((uint32_t)a) gbl_xor(w, (uint32_t)a);
    if ((uint32_t)(a >> 32)) gbl_xor(w + 1, (uint32_t)(a >> 32));
    if (gm) gbl_xor(w + 2, (uint32_t)gm);
  } else {
    uint64_t* w = reinterpret_cast<uint64_t*>(p);
    if (a) gbl_xor(w, a);
    if (gm) gbl_xor(w + 1, (uint64_t)gm);
  }
}

// ---- DSU (include/dsu.h:101-174 restated for the device) --------------------------------------
// Priorities are a bijective 32-bit mix of the vertex id, so (priority, id) is a strict total
// order and linking the lower-priority root under the higher one can never form a cycle.
__device__ inline uint32_t dsu_prio(uint32_t v, uint32_t seed) {
  uint32_t h = v ^ seed;
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}
// path halving, dsu.h:145-152, from u whose parent p is already loaded
__device__ inline uint32_t dsu_find_from(uint32_t* parent, uint32_t u, uint32_t p) {
  while (p != u) {
    const uint32_t gp = ld_acq(parent + p);
    if (gp == p) return p;
    st_rlx(parent + u, gp);
    u = gp;
    p = ld_acq(parent + u);
  }
  return u;
}
__device__ inline uint32_t dsu_find(uint32_t* parent, uint32_t u) { return dsu_find_from(parent, u, ld_acq(parent + u)); }
// find for kernels in which no thread links roots (only compression writes, which keep every
// pointer on the path to the root): cached loads are enough.
__device__ inline uint32_t dsu_find_stable(uint32_t* parent, uint32_t u) {
  uint32_t p = parent[u];
  while (true) {
    const uint32_t gp = parent[p];
    if (gp == p) return p;
    u = p;
    p = gp;
  }
}

constexpr uint32_t kNoLink = 0xFFFFFFFFu;
// CAS union ordered by rank (dsu.h:155-167): the higher hint first (when hints are given), then the
// higher priority, then the smaller id.  Any total order that stays fixed while links run
// concurrently keeps the forest acyclic (a root is only linked under a root that ranks higher), so
// the order may differ between launches.  Hints that rank high-degree vertices first let the many
// merges into a hub go through the smaller roots' own parent words instead of all contending for
// the hub's.  Returns the root that was linked under the other one, or kNoLink when a and b were
// already connected.
__device__ inline uint32_t dsu_link(uint32_t* parent, uint32_t a, uint32_t b, uint32_t pseed,
                                    const unsigned char* hint = nullptr) {
  while (true) {
    // both first parents (and hints) in one round trip; a root (the common case) needs no more
    const uint32_t a_up = ld_acq(parent + a), b_up = ld_acq(parent + b);
    uint32_t ha = hint ? hint[a] : 0, hb = hint ? hint[b] : 0;
    const uint32_t ra = dsu_find_from(parent, a, a_up), rb = dsu_find_from(parent, b, b_up);
    if (ra == rb) return kNoLink;
    if (hint && ra != a) ha = hint[ra];
    if (hint && rb != b) hb = hint[rb];
    a = ra;
    b = rb;
    const uint32_t pa = dsu_prio(a, pseed), pb = dsu_prio(b, pseed);
    const bool a_wins = ha != hb ? ha > hb : pa != pb ? pa > pb : a < b;
    if (!a_wins) { uint32_t t = a; a = b; b = t; }
    uint32_t expected = b;
    if (__hip_atomic_compare_exchange_strong(parent + b, &expected, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return b;
  }
}

// ---- spanning-forest edge set (open addressing on 64-bit edge ids) ----------------------------
constexpr uint64_t kEmpty = ~0ull;
__device__ inline uint64_t mix64(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
  return k;
}
__device__ inline void sf_insert(const DsuDev& d, uint64_t key) {
  uint64_t h = mix64(key) & d.sf_mask;
  for (uint64_t probe = 0; probe <= d.sf_mask; ++probe) {
    uint64_t expected = kEmpty;
    if (__hip_atomic_compare_exchange_strong(d.sf_hash + h, &expected, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) ||
        expected == key)
      return;
    h = (h + 1) & d.sf_mask;
  }
}
__device__ inline bool sf_contains(const DsuDev& d, uint64_t key) {
  uint64_t h = mix64(key) & d.sf_mask;
  for (uint64_t probe = 0; probe <= d.sf_mask; ++probe) {
    uint64_t cur = ld_acq(d.sf_hash + h);
    if (cur == key) return true;
    if (cur == kEmpty) return false;
    h = (h + 1) & d.sf_mask;
  }
  return false;
}

// Sample records of round 0: a GOOD sample stores its edge index, whose endpoints are both below
// n <= 2^32 - 1 after the range check, so it never equals the two sentinels (high half all ones)
// that stand for ZERO and FAIL.
constexpr uint64_t kSampZero = ~0ull;
constexpr uint64_t kSampFail = ~0ull - 1;

// An index whose endpoints are out of range (a checksum collision) counts as FAIL and raises the
// bad-sample flag (plain store: every writer stores the same value).
__device__ inline uint64_t sample_record(int code, uint64_t idx, uint32_t n, uint32_t* flags) {
  if (code == kZero) return kSampZero;
  if (code != kGood) return kSampFail;
  if ((uint32_t)(idx >> 32) >= n || (uint32_t)idx >= n) {
    flags[kFlagBad] = 1;
    return kSampFail;
  }
  return idx;
}

// ---- sample (src/sketch.cpp:94-116) over one wave --------------------------------------------
// Lane l < D holds bucket (column, depth l), lane D the deterministic bucket.  Every lane checks its
// own record with one hash, so the deterministic bucket's check runs alongside the column's.
// Returns the code; *idx gets the sampled index on GOOD.
template <class G>
__device__ inline int wave_sample(uint64_t a, G gm, uint32_t D, uint64_t chk_bf, uint64_t* idx) {
  const uint32_t lane = lane_id();
  const uint64_t mask = __ballot(lane <= D && (G)xxh3_mix(xxh3_prep(a), chk_bf) == gm);
  const uint64_t det_a = rd_lane_u64(a, (int)D);
  const G det_g = (G)rd_lane_u64((uint64_t)gm, (int)D);
  if (det_a == 0 && det_g == 0) { *idx = 0; return kZero; }
  if ((mask >> D) & 1) { *idx = det_a; return kGood; }
  const uint64_t col = mask & ((1ull << D) - 1);
  if (col) {
    *idx = rd_lane_u64(a, __builtin_ctzll(col));
    return kGood;
  }
  *idx = 0;
  return kFail;
}

// ================================================================================================
// fills
// ================================================================================================
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ================================================================================================
// Sketch update: Sketch::update over a batch (src/sketch.cpp:69-83) fused with the delta merge
// (src/cc_sketch_alg.cpp:106-118, src/sketch.cpp:149-154).
//
// One wave (a 64-thread workgroup) per tile of <= tile-size destinations of one vertex, so a CU
// keeps up to 32 tiles in flight: most tiles of a flush hold a handful of updates, and a tile's
// time is the latency of its few dependent memory round trips, not bandwidth.  Lanes take
// (update, column) pairs; the tile's delta for depths 0..kLdsDepth-1 of every column lives in LDS
// (column pitch kLdsPitch = 9, odd, so lanes on consecutive columns hit different banks), and the
// rare deeper contributions (one pair in 2^kLdsDepth) are XORed straight into the vertex's records
// with device-scope atomics (setting the column's deep flag when the record lies past the first
// line of the column's overflow run).  The deterministic bucket is kept in registers and reduced
// once per tile.  The write-back is a read-xor-write of the touched records, or device-scope
// atomic XOR when the vertex is split across tiles of the same launch.  Workgroups walk the tiles
// grid-stride (no shared counter), so a persistent grid can cover a tile count that only the
// device knows.
// ================================================================================================
constexpr uint32_t kLdsDepth = 9, kLdsPitch = 9;  // depths 0..8 per column; 9 is odd
// s_waitcnt immediate for vmcnt(0), with expcnt and lgkmcnt at their maxima (not waited on)
constexpr unsigned kWaitVmcnt0 = 0x0F70;

// WIDE: more than 28 columns, so the write-back takes more than one pass over the LDS delta (a
// separate instance keeps the one-pass form free of the loop, whose merged paths would make the
// compiler wait on this tile's stores at the next tile's first use of the prefetch register).
template <class G, bool RAW, bool WIDE>
__global__ void __launch_bounds__(kWave, 8)  // 8 waves per SIMD: 32 tiles in flight per CU
k_update_tiles(const Tile* __restrict__ tiles, const uint32_t* __restrict__ ntiles_dev, uint32_t ntiles_host,
               const uint32_t* __restrict__ dsts, const uint64_t* __restrict__ raw_idx,
               unsigned char* __restrict__ slab, Geo g, const uint64_t* __restrict__ col_bf) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t C = g.C, D = g.D;
  const uint32_t det_slot = C * kLdsPitch;
  const uint32_t lds_n = det_slot + 1;
  // per chunk: each update's pre-mixed index and checksum, read by the lanes that take its columns
  ulonglong2* s_pc = reinterpret_cast<ulonglong2*>(smem);
  uint64_t* s_bf = reinterpret_cast<uint64_t*>(smem + kWave * sizeof(ulonglong2));
  uint64_t* s_a = s_bf + C;
  G* s_g = reinterpret_cast<G*>(s_a + lds_n);
  uint32_t* s_bo = reinterpret_cast<uint32_t*>(s_g + lds_n);  // each delta entry's byte in the row
  const uint32_t lane = threadIdx.x;

  for (uint32_t i = lane; i < C; i += kWave) s_bf[i] = col_bf[i];
  // The delta is zeroed once here; afterwards the write-back clears every entry it reads.  Entries
  // at depth >= D (or at a pitch padding, if the pitch exceeds kLdsDepth) are never written:
  // offset 0.
  for (uint32_t i = lane; i < lds_n; i += kWave) {
    s_a[i] = 0;
    s_g[i] = 0;
    const uint32_t cc = i / kLdsPitch, d = i - cc * kLdsPitch;
    s_bo[i] = i == det_slot ? det_col_off(g) : (d < kLdsDepth && d < D ? rec_col_off<G>(g, cc, d) : 0u);
  }
  const uint32_t ntiles = ntiles_dev ? *ntiles_dev : ntiles_host;
  const uint64_t depth_cap = 1ull << D;
  // pair p = i * C + c: lane starts at pair `lane` and advances 64 pairs (di updates, dc columns)
  const uint32_t i0 = lane / C, c0 = lane % C;
  const uint32_t di = kWave / C, dc = kWave % C;
  __syncthreads();

  // Updates are taken in chunks of 64: lane l loads update u0 + l once and hashes its index and
  // checksum once into a 16-byte LDS slot; the chunk's (update, column) pairs then read their
  // update's slot with one LDS load, so each pair costs one column hash and no memory round trip.
  // The next tile's descriptor and each lane's first destination (or raw index) are loaded
  // before the current tile's write-back, and that chunk is hashed between the write-back's
  // record loads and its stores, where the wave waits for those loads anyway.  No tile then
  // starts with a wait, which would also wait for the previous tile's stores (the vector memory
  // counter covers stores too).
  auto load_raw = [&](const Tile& tt, uint32_t u) -> uint64_t {
    return RAW ? raw_idx[tt.begin + u] : (uint64_t)dsts[tt.begin + u];
  };
  auto to_idx = [&](const Tile& tt, uint64_t w) -> uint64_t {
    return RAW ? w : concat_pairing(tt.vertex, (uint32_t)w);
  };
  // per-lane XOR of the tile's indices and checksums: the deterministic bucket's delta
  uint64_t det_a = 0;
  G det_g = 0;
  auto hash_chunk = [&](const Tile& tt, uint32_t u0, uint64_t raw) {
    const uint32_t m = tt.count - u0 < kWave ? tt.count - u0 : kWave;
    uint64_t my_pre = 0;
    G my_chk = 0;
    if (lane < m) {
      const uint64_t idx = to_idx(tt, raw);
      my_pre = xxh3_prep(idx);
      my_chk = (G)xxh3_mix(my_pre, g.chk_bf);
      det_a ^= idx;
      det_g ^= my_chk;
    }
    s_pc[lane] = make_ulonglong2(my_pre, (uint64_t)my_chk);  // the wave's LDS ops stay in order
  };
  uint32_t t = blockIdx.x;
  Tile tl{0, 0, 0, 0};
  if (t < ntiles) {
    tl = tiles[t];
    hash_chunk(tl, 0, lane < tl.count ? load_raw(tl, lane) : 0);
  }
  while (t < ntiles) {
    unsigned char* row = slab + (size_t)tl.vertex * g.row;

    for (uint32_t u0 = 0; u0 < tl.count; u0 += kWave) {
      const uint32_t m = tl.count - u0 < kWave ? tl.count - u0 : kWave;
      if (u0 != 0) hash_chunk(tl, u0, lane < m ? load_raw(tl, u0 + lane) : 0);
      const uint32_t np = m * C;
      const uint32_t iters = (np + kWave - 1) / kWave;  // uniform over the wave
      uint32_t q = i0, c = c0;
      for (uint32_t it = 0; it < iters; ++it) {
        if (it * kWave + lane < np) {  // q < m here
          const ulonglong2 pc = s_pc[q];
          const uint64_t pre = pc.x;
          const G chk = (G)pc.y;
          const uint64_t idx = (pre << 32) | (pre >> 32);  // xxh3_prep swaps the halves
          const uint64_t h = xxh3_mix(pre, s_bf[c]);
          const uint32_t depth = (uint32_t)__builtin_ctzll(h | depth_cap);  // == D: no bucket
          if (depth < D) {
            if (depth < kLdsDepth) {
              const uint32_t li = c * kLdsPitch + depth;
              lds_xor(s_a + li, idx);
              lds_xor(s_g + li, chk);
            } else {
              xor_rec_atomic<G>(row + rec_col_off<G>(g, c, depth), idx, chk);
              if (kDeepFlags<G> && rec_is_deep(g, depth))  // the flag only ever goes from 0 to 1
                st_rlx(reinterpret_cast<uint32_t*>(row + deep_flag_off(c)), 1u);
            }
          }
        }
        q += di;
        c += dc;
        if (c >= C) {
          c -= C;
          ++q;
        }
      }
    }
    {
      const uint64_t da = wave_xor_dpp_u64(det_a);
      const G dg = sizeof(G) == 4 ? (G)wave_xor_dpp_u32((uint32_t)det_g) : (G)wave_xor_dpp_u64((uint64_t)det_g);
      if (lane == 0 && (da | (uint64_t)dg)) {
        s_a[det_slot] ^= da;
        s_g[det_slot] ^= dg;
      }
      det_a = 0;
      det_g = 0;
    }
    __syncthreads();

    const uint32_t tn = t + gridDim.x;
    // The previous tile's stores complete before this tile's loads are issued (measured: letting
    // them overlap is slower).
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    Tile tl_next{0, 0, 0, 0};
    uint64_t raw0_next = 0;
    if (tn < ntiles) {
      tl_next = tiles[tn];
      if (lane < tl_next.count) raw0_next = load_raw(tl_next, lane);
    }

    // write-back: read-xor-write of the touched records, four LDS entries per lane in flight
    auto write_back = [&](uint32_t e0, bool first) {
      uint64_t da[4], ga[4];
      G dg[4], gg[4];
      uint32_t bo[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t e = e0 + k * kWave + lane;
        da[k] = 0;
        dg[k] = 0;
        bo[k] = 0;
        if (e < lds_n) {
          bo[k] = s_bo[e];
          da[k] = s_a[e];
          dg[k] = s_g[e];
          if ((da[k] | (uint64_t)dg[k]) != 0) {
            s_a[e] = 0;
            s_g[e] = 0;
          }
        }
      }
      if (!tl.split) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((da[k] | (uint64_t)dg[k]) != 0) ld_rec<G>(row + bo[k], &ga[k], &gg[k]);
      }
      // The next tile's first chunk, hashed while this tile's record loads are in flight.  The
      // empty asm, on every path of the first pass, keeps the compiler from hoisting the hash
      // above the loads (it would wait there for the prefetch) and leaves no prefetch pending
      // into the next tile, where a reuse of its register would wait on all of this tile's
      // stores.  The record loads leave ga/gg unset for entries they skip, so no zero fill of
      // those registers waits on earlier memory operations either.
      if (first) {
        uint32_t lo = (uint32_t)raw0_next, hi = (uint32_t)(raw0_next >> 32);
        asm volatile("" : "+v"(lo), "+v"(hi)::"memory");
        if (tn < ntiles) hash_chunk(tl_next, 0, ((uint64_t)hi << 32) | lo);
      }
      if (tl.split) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((da[k] | (uint64_t)dg[k]) != 0) xor_rec_atomic<G>(row + bo[k], da[k], dg[k]);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((da[k] | (uint64_t)dg[k]) != 0) st_rec<G>(row + bo[k], ga[k] ^ da[k], (G)(gg[k] ^ dg[k]));
      }
    };
    write_back(0, true);  // the delta holds at least the deterministic bucket's entry
    for (uint32_t e0 = 4 * kWave; WIDE && e0 < lds_n; e0 += 4 * kWave) {
      write_back(e0, false);
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // none of these passes' loads stays pending into the next tile
    }
    __syncthreads();
    t = tn;
    tl = tl_next;
  }
}

// ================================================================================================
// Grouping a stream of (src,dst) pairs by vertex: the device form of the reference's gutters
// (both directions of every update, include/graph_sketch_driver.h:171-172).
// ================================================================================================
// Pairs naming a vertex >= n are skipped by every kernel and reported through bad_flag.
__global__ void k_count_pairs(const uint2* __restrict__ pairs, size_t count, uint32_t n, uint32_t* __restrict__ cnt,
                              uint32_t* __restrict__ bad_flag) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const uint2 p = pairs[i];
    if (p.x >= n || p.y >= n) {
      st_rlx(bad_flag, 1u);
      continue;
    }
    atomicAdd(cnt + p.x, 1u);
    atomicAdd(cnt + p.y, 1u);
  }
}
__global__ void k_scatter_pairs(const uint2* __restrict__ pairs, size_t count, uint32_t n,
                                uint32_t* __restrict__ cursor, uint32_t* __restrict__ dst_arr) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const uint2 p = pairs[i];
    if (p.x >= n || p.y >= n) continue;
    const uint32_t pu = atomicAdd(cursor + p.x, 1u);
    dst_arr[pu] = p.y;
    const uint32_t pv = atomicAdd(cursor + p.y, 1u);
    dst_arr[pv] = p.x;
  }
}

// Exclusive scan of per-vertex counts (and of per-vertex tile counts ceil(cnt/tile) when tile>0),
// packed as (tiles << 32 | cnt) in one u64 so a single pass scans both.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanSpan = kScanBlock * kScanItems;

__device__ inline uint64_t pack_ct(uint32_t c, uint32_t tile) {
  const uint64_t nt = tile ? (c + tile - 1) / tile : 0;
  return (nt << 32) | c;
}
__device__ inline uint64_t block_excl_scan_u64(uint64_t x, uint64_t* total) {
  __shared__ uint64_t wsum[kScanBlock / kWave];
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint64_t incl = wave_incl_scan_u64(x);
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
  for (uint32_t w = 0; w < kScanBlock / kWave; ++w) {
    if (w < wid) pre += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return pre + incl - x;
}
__global__ void __launch_bounds__(kScanBlock)
k_scan_partials(const uint32_t* __restrict__ cnt, uint32_t n, uint32_t tile, uint64_t* __restrict__ partials) {
  const uint32_t base = blockIdx.x * kScanSpan + threadIdx.x * kScanItems;
  uint64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = base + k;
    if (i < n) s += pack_ct(cnt[i], tile);
  }
  uint64_t tot;
  block_excl_scan_u64(s, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(kScanBlock)
k_scan_block_sums(uint64_t* __restrict__ partials, uint32_t nparts, uint32_t n, uint32_t* __restrict__ off,
                  uint32_t* __restrict__ toff, uint32_t* __restrict__ totals) {
  const uint32_t per = (nparts + kScanBlock - 1) / kScanBlock;
  const uint32_t b0 = threadIdx.x * per;
  uint64_t s = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < nparts) s += partials[b0 + k];
  uint64_t tot;
  uint64_t run = block_excl_scan_u64(s, &tot);
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < nparts) {
      const uint64_t v = partials[b0 + k];
      partials[b0 + k] = run;
      run += v;
    }
  if (threadIdx.x == 0) {
    totals[0] = (uint32_t)tot;
    totals[1] = (uint32_t)(tot >> 32);
    off[n] = (uint32_t)tot;
    if (toff) toff[n] = (uint32_t)(tot >> 32);
  }
}
__global__ void __launch_bounds__(kScanBlock)
k_scan_apply(const uint32_t* __restrict__ cnt, uint32_t n, uint32_t tile, const uint64_t* __restrict__ partials,
             uint32_t* __restrict__ off, uint32_t* __restrict__ toff, uint32_t* __restrict__ cursor) {
  const uint32_t base = blockIdx.x * kScanSpan + threadIdx.x * kScanItems;
  uint32_t c[kScanItems];
  uint64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = base + k;
    c[k] = i < n ? cnt[i] : 0;
    s += pack_ct(c[k], tile);
  }
  uint64_t tot;
  uint64_t run = block_excl_scan_u64(s, &tot) + partials[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = base + k;
    if (i < n) {
      off[i] = (uint32_t)run;
      if (toff) toff[i] = (uint32_t)(run >> 32);
      if (cursor) cursor[i] = (uint32_t)run;
    }
    run += pack_ct(c[k], tile);
  }
}
// Also accumulates, when acc is given, the flush's algorithmic bytes (sum over vertices of
// 4k + 2*min(S, k*(C+1)*B)) into acc[blockIdx.x]: one slot per block, so no atomics; the grid
// never exceeds kProfSlots blocks and launches on one stream are ordered.
__global__ void __launch_bounds__(256)
k_make_tiles(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off, const uint32_t* __restrict__ toff,
             uint32_t n, uint32_t tile, Tile* __restrict__ tiles, uint64_t* __restrict__ acc, uint64_t S,
             uint64_t per_update) {
  uint64_t bytes = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t c = cnt[v];
    if (!c) continue;
    const uint64_t touched = (uint64_t)c * per_update;
    bytes += 4ull * c + 2 * (touched < S ? touched : S);
    const uint32_t nt = (c + tile - 1) / tile, o = off[v], t0 = toff[v];
    for (uint32_t j = 0; j < nt; ++j) {
      const uint32_t b = j * tile;
      tiles[t0 + j] = Tile{v, o + b, (c - b) < tile ? (c - b) : tile, nt > 1 ? 1u : 0u};
    }
  }
  if (!acc) return;  // uniform over the grid
  __shared__ uint64_t s_part[256 / kWave];
  for (int o = 32; o > 0; o >>= 1) bytes += shfl_xor_u64(bytes, o);
  if (lane_id() == 0) s_part[threadIdx.x / kWave] = bytes;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < 256 / kWave; ++w) t += s_part[w];
    acc[blockIdx.x] += t;
  }
}

// Grouping by vertex without device-scope atomics (vertex counts up to 2^24): a bucket partition
// over vertex ranges of R = 2^shift vertices (at most kPartBins ranges).
//   1. k_part_hist: each workgroup takes a contiguous chunk of the pairs and counts the directed
//      entries per range in LDS; the counts go to hist[range * nblk + block].
//   2. an exclusive scan of hist (range-major), so each (range, block) owns a slot interval.
//   3. k_part_scatter: the same chunks again; LDS cursors place the records (vertex, other end)
//      in range order.
//   4. k_part_group: one workgroup per range counts its vertices' entries in LDS, scans them, and
//      writes cnt[v] and each vertex's destinations contiguously (vertex order, the same layout
//      the vertex-count scan then gives).  The order of one vertex's destinations is arbitrary,
//      which XOR does not see.
constexpr uint32_t kPartBins = 4096, kPartBlock = 1024, kPartMaxBlocks = 256, kPartMaxRange = 4096;

__device__ inline void part_chunk(size_t count, uint32_t nblk, size_t* lo, size_t* hi) {
  const size_t per = (count + nblk - 1) / nblk;
  *lo = (size_t)blockIdx.x * per;
  *hi = *lo + per < count ? *lo + per : count;
}

__global__ void __launch_bounds__(kPartBlock)
k_part_hist(const uint2* __restrict__ pairs, size_t count, uint32_t n, uint32_t shift, uint32_t bins, uint32_t nblk,
            uint32_t* __restrict__ hist, uint32_t* __restrict__ bad_flag) {
  __shared__ uint32_t h[kPartBins];
  for (uint32_t b = threadIdx.x; b < bins; b += kPartBlock) h[b] = 0;
  __syncthreads();
  size_t lo, hi;
  part_chunk(count, nblk, &lo, &hi);
  for (size_t k = lo + threadIdx.x; k < hi; k += kPartBlock) {
    const uint2 p = pairs[k];
    if (p.x >= n || p.y >= n) {
      st_rlx(bad_flag, 1u);
      continue;
    }
    atomicAdd(h + (p.x >> shift), 1u);
    atomicAdd(h + (p.y >> shift), 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < bins; b += kPartBlock) hist[(size_t)b * nblk + blockIdx.x] = h[b];
}

__global__ void __launch_bounds__(kPartBlock)
k_part_scatter(const uint2* __restrict__ pairs, size_t count, uint32_t n, uint32_t shift, uint32_t bins,
               uint32_t nblk, const uint32_t* __restrict__ hoff, uint2* __restrict__ rec) {
  __shared__ uint32_t cur[kPartBins];
  for (uint32_t b = threadIdx.x; b < bins; b += kPartBlock) cur[b] = hoff[(size_t)b * nblk + blockIdx.x];
  __syncthreads();
  size_t lo, hi;
  part_chunk(count, nblk, &lo, &hi);
  for (size_t k = lo + threadIdx.x; k < hi; k += kPartBlock) {
    const uint2 p = pairs[k];
    if (p.x >= n || p.y >= n) continue;
    rec[atomicAdd(cur + (p.x >> shift), 1u)] = make_uint2(p.x, p.y);
    rec[atomicAdd(cur + (p.y >> shift), 1u)] = make_uint2(p.y, p.x);
  }
}

__global__ void __launch_bounds__(kPartBlock)
k_part_group(const uint2* __restrict__ rec, uint32_t n, uint32_t shift, uint32_t bins, uint32_t nblk,
             const uint32_t* __restrict__ hoff, uint32_t* __restrict__ cnt, uint32_t* __restrict__ dst_arr) {
  __shared__ uint32_t lc[kPartMaxRange];
  __shared__ uint32_t wsum[kPartBlock / kWave];
  const uint32_t b = blockIdx.x;
  const uint32_t base = b << shift, R = 1u << shift;
  const uint32_t nv = n - base < R ? n - base : R;
  const uint32_t s = hoff[(size_t)b * nblk];
  const uint32_t e = hoff[(size_t)(b + 1) * nblk];  // hoff[bins * nblk] holds the total
  for (uint32_t k = threadIdx.x; k < R; k += kPartBlock) lc[k] = 0;
  __syncthreads();
  for (uint32_t j = s + threadIdx.x; j < e; j += kPartBlock) atomicAdd(lc + (rec[j].x - base), 1u);
  __syncthreads();
  // exclusive scan of lc[0..R): each thread owns a run of R / kPartBlock (>= 1) consecutive slots
  const uint32_t per = (R + kPartBlock - 1) / kPartBlock;
  const uint32_t k0 = threadIdx.x * per;
  uint32_t mine = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (k0 + k < R) {
      const uint32_t c = lc[k0 + k];
      if (k0 + k < nv) cnt[base + k0 + k] = c;
      mine += c;
    }
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint64_t incl = wave_incl_scan_u64(mine);
  if (lane == kWave - 1) wsum[wid] = (uint32_t)incl;
  __syncthreads();
  uint32_t run = (uint32_t)incl - mine;
  for (uint32_t w = 0; w < wid; ++w) run += wsum[w];
  for (uint32_t k = 0; k < per; ++k)
    if (k0 + k < R) {
      const uint32_t c = lc[k0 + k];
      lc[k0 + k] = s + run;  // becomes the vertex's cursor
      run += c;
    }
  __syncthreads();
  for (uint32_t j = s + threadIdx.x; j < e; j += kPartBlock) {
    const uint2 r = rec[j];
    dst_arr[atomicAdd(lc + (r.x - base), 1u)] = r.y;
  }
}

// The apply step of the scan when there are at most kScanBlock partials: every block adds up the
// partials before its own (one per thread), so the separate block-sums launch is not needed, and
// the last block writes the totals.  tiles != nullptr: each block also writes its vertices' tiles
// (what k_make_tiles does), with the profiling bytes when acc != nullptr.
__global__ void __launch_bounds__(kScanBlock)
k_scan_apply_direct(const uint32_t* __restrict__ cnt, uint32_t n, uint32_t tile, const uint64_t* __restrict__ partials,
                    uint32_t nparts, uint32_t* __restrict__ off, uint32_t* __restrict__ toff,
                    uint32_t* __restrict__ cursor, uint32_t* __restrict__ totals, Tile* __restrict__ tiles,
                    uint64_t* __restrict__ acc, uint64_t S, uint64_t per_update) {
  uint64_t prefix;
  block_excl_scan_u64(threadIdx.x < blockIdx.x ? partials[threadIdx.x] : 0, &prefix);
  const uint32_t base = blockIdx.x * kScanSpan + threadIdx.x * kScanItems;
  uint32_t c[kScanItems];
  uint64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = base + k;
    c[k] = i < n ? cnt[i] : 0;
    s += pack_ct(c[k], tile);
  }
  uint64_t tot;
  uint64_t run = block_excl_scan_u64(s, &tot) + prefix;
  uint64_t bytes = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = base + k;
    if (i < n) {
      off[i] = (uint32_t)run;
      if (toff) toff[i] = (uint32_t)(run >> 32);
      if (cursor) cursor[i] = (uint32_t)run;
      if (tiles && c[k]) {
        const uint32_t nt = (c[k] + tile - 1) / tile, o = (uint32_t)run, t0 = (uint32_t)(run >> 32);
        for (uint32_t j = 0; j < nt; ++j) {
          const uint32_t b = j * tile;
          tiles[t0 + j] = Tile{i, o + b, (c[k] - b) < tile ? (c[k] - b) : tile, nt > 1 ? 1u : 0u};
        }
        const uint64_t touched = (uint64_t)c[k] * per_update;
        bytes += 4ull * c[k] + 2 * (touched < S ? touched : S);
      }
    }
    run += pack_ct(c[k], tile);
  }
  if (blockIdx.x == nparts - 1 && threadIdx.x == 0) {
    const uint64_t all = prefix + tot;
    totals[0] = (uint32_t)all;
    totals[1] = (uint32_t)(all >> 32);
    off[n] = (uint32_t)all;
    if (toff) toff[n] = (uint32_t)(all >> 32);
  }
  if (acc) {  // profiling: the block's algorithmic bytes
    for (int o = 32; o > 0; o >>= 1) bytes += shfl_xor_u64(bytes, o);
    if (lane_id() == 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(acc + blockIdx.x % kProfSlots), (unsigned long long)bytes);
  }
}

// Returns whether the tile list was written too (only when tiles != nullptr and the one-pass apply
// was used).
bool scan_counts(const uint32_t* cnt, uint32_t n, uint32_t tile, uint64_t* partials, uint32_t* off, uint32_t* toff,
                 uint32_t* cursor, uint32_t* totals, hipStream_t s, Tile* tiles = nullptr,
                 const FlushBytes* fb = nullptr) {
  const uint32_t nparts = (n + kScanSpan - 1) / kScanSpan;
  hipLaunchKernelGGL(k_scan_partials, dim3(nparts), dim3(kScanBlock), 0, s, cnt, n, tile, partials);
  if (nparts <= kScanBlock) {
    hipLaunchKernelGGL(k_scan_apply_direct, dim3(nparts), dim3(kScanBlock), 0, s, cnt, n, tile, partials, nparts, off,
                       toff, cursor, totals, tiles, fb ? fb->acc : nullptr, fb ? fb->S : 0, fb ? fb->per_update : 0);
    return tiles != nullptr;
  }
  hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(kScanBlock), 0, s, partials, nparts, n, off, toff, totals);
  hipLaunchKernelGGL(k_scan_apply, dim3(nparts), dim3(kScanBlock), 0, s, cnt, n, tile, partials, off, toff, cursor);
  return false;
}

inline uint32_t grid_for(size_t work, uint32_t block, uint32_t cap = 8192) {
  size_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)(g < cap ? g : cap);
}

// ================================================================================================
// Eager DSU over a batch of updates (pre_insert, src/cc_sketch_alg.cpp:79-104), in three phases
// separated by kernel boundaries:
//   1. every update attempts the CAS union of its endpoints; winners are marked merged;
//   2. merged edges are inserted into the spanning-forest set;
//   3. an update that did not merge and whose edge is in the forest invalidates the DSU.
// The outcome equals the reference's for the linearisation the CAS order defines (the
// reference's stream threads are concurrent as well): a deleted forest edge always invalidates,
// and any edge that merged and occurs again in the batch invalidates.
// ================================================================================================
// src != pairs: the batch is also copied from src into pairs (the staging buffer) on the way, so
// ingesting device pairs needs no separate copy launch; the copy happens whether or not the DSU
// is still valid.
// build_sf: the first batch after a query also fills the forest membership set from the query's
// forest edges (sf_edge) -- lock-free inserts like this batch's own, so the two need no order.
__global__ void k_eager_union(const uint2* __restrict__ src, uint2* __restrict__ pairs, size_t count,
                              uint8_t* __restrict__ merged, DsuDev d, uint32_t build_sf) {
  const bool live = !ld_acq(d.flags + kFlagEagerInvalid);
  if (!live && src == pairs) return;
  if (build_sf && live)
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < d.n; v += gridDim.x * blockDim.x) {
      const uint64_t e = d.sf_edge[v];
      if (e != kEmpty) sf_insert(d, e);
    }
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const uint2 p = src[i];
    if (src != pairs) pairs[i] = p;
    if (!live) continue;
    if (p.x >= d.n || p.y >= d.n) {
      merged[i] = 0;
      st_rlx(d.flags + kFlagBadInput, 1u);
      continue;
    }
    const uint32_t child = dsu_link(d.parent, p.x, p.y, d.prio_seed);
    merged[i] = child != kNoLink ? 1 : 0;
    if (child != kNoLink) {
      const uint64_t key = concat_pairing(p.x, p.y);
      d.sf_edge[child] = key;
      sf_insert(d, key);  // into the forest set; the deletion check runs once the whole batch is merged
    }
  }
}
__global__ void k_eager_sf_check(const uint2* __restrict__ pairs, size_t count, const uint8_t* __restrict__ merged,
                                 DsuDev d) {
  if (ld_acq(d.flags + kFlagEagerInvalid)) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    if (merged[i]) continue;
    const uint2 p = pairs[i];
    if (p.x == p.y || p.x >= d.n || p.y >= d.n) continue;
    if (sf_contains(d, concat_pairing(p.x, p.y))) {
      d.flags[kFlagEagerInvalid] = 1;  // same value from every writer
      if (d.inv_mirror) {  // lets the host stop launching eager batches without a synchronisation
        *d.inv_mirror = 1;
        __threadfence_system();
      }
    }
  }
}

// ================================================================================================
// Borůvka (boruvka_emulation, src/cc_sketch_alg.cpp:464-513)
// ================================================================================================


// Round 0 (run_round_zero, 223-244): every vertex samples column 0 of its own sketch
// (Sketch::sample, src/sketch.cpp:94-116).  A wave takes four vertices in groups of 16 lanes: lane
// 16j + i (i < 15) checks depth i of vertex v0 + j and lane 16j + 15 its deterministic bucket, so
// one hash per lane serves four vertices.  The first good depth of a vertex is almost always below
// 15 (a bucket at depth d holds about deg / 2^(d+1) edges); a vertex with no good bucket there
// (very high degree, or none at all) is finished by the whole wave over depths 15..D-1.
template <class G>
__global__ void __launch_bounds__(256)
k_sample_round0(const unsigned char* __restrict__ slab, Geo g, uint64_t* __restrict__ samp,
                uint32_t* __restrict__ flags, unsigned char* __restrict__ det_cpy, unsigned char* __restrict__ hint,
                DsuDev d) {
  constexpr uint32_t V = 4, GL = kWave / V, GD = GL - 1;  // vertices per wave, lanes per vertex, depths
  constexpr uint32_t RB = 8 + sizeof(G);
  // Round-0 hooking: every vertex is a root when round 0 begins, and only its own thread would link
  // it under a higher-ranked vertex, so a sampled edge (v, u) whose other end u ranks above v
  // (priority, then the smaller id) is merged here with plain stores -- v under u, parents always
  // ranking higher, so the forest stays acyclic.  The other samples wait in samp for
  // k_apply_samples, which links them with the CAS union.
  bool hooked = false;
  auto put = [&](uint32_t v, uint64_t rec) {
    if (rec < kSampFail) {  // a GOOD sample: an edge index with both ends below n
      const uint32_t x = (uint32_t)(rec >> 32), y = (uint32_t)rec;
      const uint32_t u = x == v ? y : x;
      if ((x == v || y == v) && u != v) {
        const uint32_t pv = dsu_prio(v, d.prio_seed), pu = dsu_prio(u, d.prio_seed);
        if (pu > pv || (pu == pv && u < v)) {
          d.parent[v] = u;
          d.sf_edge[v] = rec;  // the forest edge, in the slot of the root it linked
          samp[v] = kSampZero;
          hooked = true;
          return;
        }
      }
    }
    samp[v] = rec;
  };
  if (ld_acq(d.flags + kFlagSkip)) return;  // the query keeps the eager DSU
  const uint32_t lane = lane_id(), i = lane % GL, base = lane - i;
  const uint32_t nd = g.D < GD ? g.D : GD;  // depths checked in the grouped pass
  const uint32_t waves = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave * V; v0 < g.n; v0 += waves * V) {
    const uint32_t v = v0 + lane / GL;
    uint64_t a = 0;
    G gm = 0;
    const bool live = v < g.n && (i < nd || i == GD);
    if (live) ld_rec<G>(slab + (size_t)v * g.row + (i == GD ? det_col_off(g) : rec_col_off<G>(g, 0, i)), &a, &gm);
    const uint64_t mask = __ballot(live && (G)xxh3_mix(xxh3_prep(a), g.chk_bf) == gm);
    const uint32_t gmask = (uint32_t)(mask >> base) & ((1u << GL) - 1);
    const uint32_t colm = gmask & ((1u << nd) - 1);
    // degree class: one past the deepest nonzero bucket among the grouped depths (a bucket at depth
    // d is nonzero about when the degree reaches 2^d)
    const uint64_t nzm = __ballot(live && i < nd && (a | (uint64_t)gm) != 0);
    const uint32_t gnz = (uint32_t)(nzm >> base) & ((1u << nd) - 1);
    // every lane takes part in the shuffles; lane i == 0 of a group decides
    const uint64_t det_a = shfl_u64(a, (int)(base + GD));
    const G det_g = (G)shfl_u64((uint64_t)gm, (int)(base + GD));
    const uint64_t col_a = shfl_u64(a, (int)(base + (colm ? __builtin_ctz(colm) : 0)));
    bool undecided = false;
    if (i == 0 && v < g.n) {
      hint[v] = (unsigned char)(gnz ? 32 - __builtin_clz(gnz) : 0);
      if (det_a == 0 && det_g == 0) samp[v] = kSampZero;
      else if ((gmask >> GD) & 1) put(v, sample_record(kGood, det_a, g.n, flags));
      else if (colm) put(v, sample_record(kGood, col_a, g.n, flags));
      else if (g.D <= GD) samp[v] = kSampFail;
      else undecided = true;
    }
    // the deterministic bucket does not depend on the round: later rounds read this compact copy
    if (i == GD && v < g.n) st_rec<G>(det_cpy + (size_t)v * RB, a, gm);
    for (uint64_t und = __ballot(undecided); und; und &= und - 1) {  // uniform over the wave
      const uint32_t u = v0 + (uint32_t)__builtin_ctzll(und) / GL;
      const uint32_t dep = GD + lane;  // D - GD <= 64 lanes for every supported n
      uint64_t b = 0;
      G bg = 0;
      if (dep < g.D) ld_rec<G>(slab + (size_t)u * g.row + rec_col_off<G>(g, 0, dep), &b, &bg);
      const uint64_t m2 = __ballot(dep < g.D && (G)xxh3_mix(xxh3_prep(b), g.chk_bf) == bg);
      const uint64_t ia = m2 ? rd_lane_u64(b, __builtin_ctzll(m2)) : 0;
      if (lane == 0) put(u, m2 ? sample_record(kGood, ia, g.n, flags) : kSampFail);
    }
  }
  if (__ballot(hooked) && lane_id() == 0) d.mod[0] = 1;  // round 0 merged something
}

// The merges of a round (sample_supernode, src/cc_sketch_alg.cpp:137-166), one thread per
// supernode: every vertex that was a root when the round began (root == nullptr: every vertex, as
// in round 0) acts on its sample record.  The round is modified by a FAIL or a merge.  Links rank
// roots by the degree class round 0 recorded (hint), so hubs stay roots and absorb their
// neighbours without contention.
constexpr uint32_t kNoRoot = 0xFFFFFFFFu;  // no supernode (unused partial slot, no giant)

// Round state words (RoundBufs::totals).  A query starts with kStGiantFirst = kNoRoot.
//   kStSeg + 2(r & 1) entries regrouped in round r (the level-1 entry count, and the length of the
//                     active list of round r + 1)
//   kStGiantRound     root of this round's giant supernode, or kNoRoot
//   kStGiantFirst     the root elected giant first in this query (any member identifies the giant,
//                     which only grows), or kNoRoot
//   kStActMode        1 when this round walks the active list instead of every vertex
// Round 1 walks every vertex in id order.  Round r >= 2 walks the member list regrouped in round
// r - 1: every vertex outside the giant (all of them before one is elected), grouped by the
// previous round's roots, so the members of a new supernode mostly sit next to each other and a
// wave meets few distinct roots.  Vertices that joined the giant are dropped as they are met.
//   kStBest (u64)     before any giant: max over this round's roots holding >= n/kCandidateDiv
//                     vertices of (count << 32 | root); the next round excludes that supernode.
//                     The complement sum needs no majority -- any supernode's sum is the XOR of the
//                     vertices outside it -- so a large one can be left out a round early.
//   kStSeg + 2(r & 1) + 1  supernodes outside the giant in round r (entries of RoundBufs::segs[r & 1])
// The two counts of one round parity share a 64-bit word, so a block takes its share of both with
// one atomic: entries at kStSeg + 2 * (r & 1), supernodes at kStSeg + 2 * (r & 1) + 1.
constexpr uint32_t kStGiantRound = 2, kStGiantFirst = 3, kStActMode = 4, kStBest = 6, kStSeg = 8;
constexpr uint32_t kStWords = 12;
constexpr uint32_t kCandidateDiv = 32;
constexpr uint32_t kMaxStatRounds = 64;

// cnt / st (a query's start, or nullptr): the root counts are cleared too, and the round state
// words are set to their start values (no giant), so a query begins with one launch.
//
// force (a query's start, or nullptr): when it reads 0 and the eager DSU is still valid, the query
// keeps that DSU -- its answer is the cached one -- so the kernel only clears the round flags (every
// round kernel of the query then returns at once), sets the round state words to their start values
// and sets kFlagSkip.  The eager-invalid flag is only read here (every workgroup sees the same value)
// and is cleared by round 0 of a query that runs.
__global__ void k_dsu_reset(uint32_t* parent, uint32_t n, uint64_t* sf_hash, uint64_t sf_cap, uint64_t* sf_edge,
                            uint32_t* flags, uint32_t* mod, uint32_t nmod, uint32_t* cnt, uint32_t* st,
                            const uint32_t* force) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (force && !*force && !ld_acq(flags + kFlagEagerInvalid)) {
    if (t == 0) flags[kFlagSkip] = 1;
    for (size_t i = t; i < nmod; i += stride) mod[i] = 0;
    // The round kernels read their walk counts and member lists before they look at the round
    // flag, so the state words get their start values (empty walks) here as well.
    if (st && t < kStWords) st[t] = (t == kStGiantRound || t == kStGiantFirst) ? kNoRoot : 0u;
    return;
  }
  for (size_t i = t; i < n; i += stride) {
    parent[i] = (uint32_t)i;
    sf_edge[i] = kEmpty;
    if (cnt) cnt[i] = 0;
  }
  for (size_t i = t; i < sf_cap; i += stride) sf_hash[i] = kEmpty;
  // The bad-input flag is left alone: pairs grouped by the flush just before a query set it, and the
  // query's answer reports it (it is cleared where it is reported).  A query's reset leaves the
  // eager-invalid flag to round 0 (see above).
  if (t < kNumFlags && t != kFlagBadInput && !(force && t == kFlagEagerInvalid)) flags[t] = 0;
  for (size_t i = t; i < nmod; i += stride) mod[i] = 0;
  if (st && t < kStWords) st[t] = (t == kStGiantRound || t == kStGiantFirst) ? kNoRoot : 0u;
}

__device__ inline uint32_t walk_count(const uint32_t* st, uint32_t round, uint32_t n) {
  return round >= 2 ? st[kStSeg + 2 * ((round - 1) & 1)] : n;
}
__device__ inline uint32_t walk_vertex(const uint32_t* act_list, uint32_t round, uint32_t i) {
  return round >= 2 ? act_list[i] : i;
}

__device__ inline bool apply_sample(const DsuDev& d, uint64_t s, const unsigned char* hint = nullptr) {
  if (s == kSampZero) return false;
  if (s == kSampFail) return true;
  const uint32_t a = (uint32_t)(s >> 32), b = (uint32_t)s;
  const uint32_t child = dsu_link(d.parent, a, b, d.prio_seed, hint);
  if (child == kNoLink) return false;
  // A merge records its edge in the spanning forest (cc_sketch_alg.cpp:152-162), in the slot of
  // the root it linked (every root is linked at most once between resets, so slots never
  // collide); the membership set the eager DSU uses is filled from these slots by the first eager
  // batch after the query (k_eager_union), off the merge's critical path.
  d.sf_edge[child] = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
  return true;
}

// st == nullptr (round 0): every vertex is a root.  Otherwise the round's supernodes: its segment
// list (every root outside the giant) and the giant's root.
__global__ void __launch_bounds__(256)
k_apply_samples(const uint64_t* __restrict__ samp, const RoundSeg* __restrict__ segs, DsuDev d,
                const uint32_t* __restrict__ st, uint32_t round, const uint32_t* __restrict__ prev_mod,
                uint32_t* __restrict__ cur_mod, const unsigned char* __restrict__ hint) {
  uint32_t live = prev_mod ? ld_acq(prev_mod) : 1u;  // checked after the first reads
  if (!st) {  // round 0: nothing to do when the query keeps the eager DSU; else the DSU restarts valid
    if (ld_acq(d.flags + kFlagSkip)) live = 0;
    else if (blockIdx.x == 0 && threadIdx.x == 0) d.flags[kFlagEagerInvalid] = 0;
  }
  const bool act = st && st[kStActMode] != 0;
  const uint32_t gr = act ? st[kStGiantRound] : kNoRoot;
  const uint32_t count = st ? st[kStSeg + 2 * (round & 1) + 1] : d.n;
  if (!live) return;
  bool modified = false;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t v = st ? segs[i].root : i;
    if (v == gr) continue;  // handled once below
    if (apply_sample(d, samp[v], hint)) modified = true;
  }
  if (gr != kNoRoot && blockIdx.x == 0 && threadIdx.x == 0 && apply_sample(d, samp[gr], hint)) modified = true;
  if (__syncthreads_or(modified) && threadIdx.x == 0) *cur_mod = 1;  // one plain store per block
}

// ---- rounds >= 1 (perform_boruvka_round 246-380 + create_merge_instructions 382-462) ----------
// Every kernel of round r first reads mod[r-1]: when round r-1 changed nothing the whole round is
// a no-op, which lets the host queue several rounds before it looks at the flags.
//
// Regrouping by root (round 1, every vertex): A) root of every vertex, members per root, and each
// vertex's place in its root's segment (the value its root's counter had when the vertex was
// counted); B) each root gets a contiguous segment [off, off+cnt) from one running cursor
// (contiguous, not in root order, which Borůvka does not need), and the round's segment list;
// C) every vertex to off[root] + its place.  Counts are aggregated per 1024-thread block in an LDS
// table first, so a root shared by the whole block costs one device atomic per block instead of
// one per wave or per vertex.
constexpr int kAggBlock = 1024;
constexpr uint32_t kAggTable = 2048;  // >= 2 * kAggBlock, power of two
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

__device__ inline uint32_t lds_table_slot(uint32_t* keys, uint32_t key) {
  uint32_t h = (key * 2654435761u) & (kAggTable - 1);
  while (true) {
    const uint32_t prev = atomicCAS(keys + h, kEmptyKey, key);
    if (prev == kEmptyKey || prev == key) return h;
    h = (h + 1) & (kAggTable - 1);
  }
}

__global__ void __launch_bounds__(kAggBlock)
k_bv_find_count(uint32_t* __restrict__ parent, uint32_t n, uint2* __restrict__ nd, uint32_t* __restrict__ cnt,
                uint32_t* __restrict__ st, uint32_t round, const uint32_t* __restrict__ act_list,
                const uint32_t* __restrict__ prev_mod) {
  // The round flag is loaded together with the first reads (all in bounds whether or not the
  // round runs: counts and lists are those of an earlier round of this query), saving a round trip.
  const uint32_t live = ld_acq(prev_mod);
  uint32_t gfirst = st[kStGiantFirst];
  const uint64_t best = *reinterpret_cast<const uint64_t*>(st + kStBest);
  const uint32_t i = blockIdx.x * kAggBlock + threadIdx.x;
  const uint32_t walk = walk_count(st, round, n);
  const bool in_walk = i < walk;
  const uint32_t v = in_walk ? walk_vertex(act_list, round, i) : 0;
  // the grid covers n; blocks past the walk have nothing to count (block 0 writes the round state)
  if (!live || (blockIdx.x > 0 && blockIdx.x * kAggBlock >= walk)) return;
  __shared__ uint32_t t_key[kAggTable], t_cnt[kAggTable];
  for (uint32_t k = threadIdx.x; k < kAggTable; k += kAggBlock) {
    t_key[k] = kEmptyKey;
    t_cnt[k] = 0;
  }
  if (gfirst == kNoRoot && best) gfirst = (uint32_t)best;  // a candidate recorded by the previous round
  const bool act = gfirst != kNoRoot;
  const uint32_t gcur = act ? dsu_find_stable(parent, gfirst) : kNoRoot;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[kStSeg + 2 * (round & 1)] = 0;
    st[kStSeg + 2 * (round & 1) + 1] = 0;
    st[kStGiantRound] = gcur;
    st[kStGiantFirst] = gfirst;
    st[kStActMode] = act ? 1u : 0u;
  }
  __syncthreads();
  __shared__ uint32_t t_base[kAggTable];
  uint32_t r = 0;
  if (in_walk) r = dsu_find_stable(parent, v);
  __syncthreads();
  // full path compression: every walked vertex now points at its root (no links in this kernel)
  if (in_walk) parent[v] = r;
  const bool active = in_walk && r != gcur;  // members of the giant are not counted
  // every lane inserts its root at once: LDS atomics absorb lanes that share a root, and no wave
  // walks its distinct roots one by one; the returned counts give each vertex its place in its
  // root's segment
  uint32_t my_slot = 0, my_pos = 0;
  if (active) {
    my_slot = lds_table_slot(t_key, r);
    my_pos = atomicAdd(t_cnt + my_slot, 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kAggTable; k += kAggBlock)
    if (t_key[k] != kEmptyKey) t_base[k] = atomicAdd(cnt + t_key[k], t_cnt[k]);
  __syncthreads();
  if (in_walk) nd[v] = make_uint2(r, active ? t_base[my_slot] + my_pos : 0u);
}

// Segments for the roots met in this round's walk (every non-giant root is a member of its own
// supernode, so it is walked).  cnt is cleared here, its last reader.  Without a giant yet, a root
// holding more than half of the vertices is elected: at most one can.
// A root's segment goes into the round's segment list too (block scan of the counts and of the
// roots, packed in one 64-bit scan: counts in the low 40 bits, roots above).
constexpr uint32_t kSegShift = 40;

__global__ void __launch_bounds__(kAggBlock)
k_bv_alloc(uint32_t* __restrict__ cnt, uint32_t n, uint32_t* __restrict__ off, uint32_t* __restrict__ cursor,
           uint32_t* __restrict__ st, uint32_t round, const uint32_t* __restrict__ act_list, uint32_t giant_ok,
           const uint32_t* __restrict__ prev_mod, RoundSeg* __restrict__ segs) {
  const uint32_t live = ld_acq(prev_mod);  // checked after the first reads (see k_bv_find_count)
  __shared__ uint64_t wsum[kAggBlock / kWave];
  __shared__ uint32_t s_base, s_sbase;
  const bool act = st[kStActMode] != 0;
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint32_t i = blockIdx.x * kAggBlock + threadIdx.x;
  const uint32_t walk = walk_count(st, round, n);
  const bool in_walk = i < walk;
  const uint32_t v = in_walk ? walk_vertex(act_list, round, i) : 0;
  uint32_t c = in_walk ? cnt[v] : 0;
  if (!live || blockIdx.x * kAggBlock >= walk) return;  // whole blocks past the walk
  if (c) cnt[v] = 0;
  if (!act && giant_ok && 2ull * c > n) {
    st[kStGiantFirst] = v;
    st[kStGiantRound] = v;
    c = 0;
  } else if (!act && giant_ok && c >= 2 && c >= n / kCandidateDiv) {  // at most kCandidateDiv roots
    atomicMax(reinterpret_cast<unsigned long long*>(st + kStBest), ((unsigned long long)c << 32) | v);
  }
  const uint64_t x = c ? ((uint64_t)1 << kSegShift) | c : 0;
  const uint64_t incl = wave_incl_scan_u64(x);
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
  for (uint32_t w = 0; w < kAggBlock / kWave; ++w) {
    if (w < wid) pre += wsum[w];
    tot += wsum[w];
  }
  constexpr uint64_t kCntMask = ((uint64_t)1 << kSegShift) - 1;
  if (threadIdx.x == 0) {
    // entries and supernodes of the block in one 64-bit atomic (neither sum can reach 2^32)
    const uint64_t add = ((tot >> kSegShift) << 32) | (tot & kCntMask);
    const uint64_t old =
        add ? atomicAdd(reinterpret_cast<unsigned long long*>(st + kStSeg + 2 * (round & 1)), (unsigned long long)add) : 0;
    s_base = (uint32_t)old;
    s_sbase = (uint32_t)(old >> 32);
  }
  __syncthreads();
  if (c) {
    const uint64_t ex = pre + incl - x;
    const uint32_t o = s_base + (uint32_t)(ex & kCntMask);
    const uint32_t k = s_sbase + (uint32_t)(ex >> kSegShift);
    off[v] = o;
    cursor[v] = o + c;  // the segment end
    if (k < n) segs[k] = RoundSeg{v, o, c, 0};
  }
}

__global__ void __launch_bounds__(kAggBlock)
k_bv_scatter(const uint2* __restrict__ nd, uint32_t n, const uint32_t* __restrict__ off,
             uint32_t* __restrict__ members, uint32_t* __restrict__ mroot, const uint32_t* __restrict__ st, uint32_t round,
             const uint32_t* __restrict__ act_list, const uint32_t* __restrict__ prev_mod) {
  const uint32_t live = ld_acq(prev_mod);  // checked after the first reads (see k_bv_find_count)
  const uint32_t gr = st[kStGiantRound];
  const uint32_t e = st[kStSeg + 2 * (round & 1)];
  const uint32_t i = blockIdx.x * kAggBlock + threadIdx.x;
  const uint32_t walk = walk_count(st, round, n);
  const bool in_walk = i < walk;
  const uint32_t v = in_walk ? walk_vertex(act_list, round, i) : 0;
  const uint2 t = in_walk && v < n ? nd[v] : make_uint2(0u, 0u);
  if (!live || !in_walk || v >= n) return;
  if (t.x == gr || t.x >= n) return;  // the giant supernode's members are not regrouped
  const uint32_t p = off[t.x] + t.y;
  if (p >= e || p >= n) return;
  members[p] = v;  // the root beside the member, so the reduce reads both in list order
  mroot[p] = t.x;
}

// ---- regrouping by segments (rounds >= 2) ------------------------------------------------------
// A supernode of round r is a union of supernodes of round r - 1, and those are contiguous runs
// ("segments") of round r - 1's member list.  So round r regroups whole runs: A) the new root of
// every old segment, and the run's offset inside its new root's segment, taken from the root's
// member counter (atomics aggregated per block in LDS); B) a contiguous segment per new root and
// the round's segment list; C) one pass over the old member list copying each member to its place
// (its root's segment start + its run's offset + its place in the run), with its new root beside
// it and its parent pointer compressed to that root.  A and B walk the old segments (one per old
// supernode), only C walks the members, in list order.  Runs whose new root is the giant are dropped, as the vertex walk drops
// the giant's members.  Every kernel loads its counts and lists before it looks at the round flag
// (all in range: they are those of an earlier round of this query, or the start values) and writes
// nothing when the round is not live; indices computed from device data are bounded by n before
// every store.

__global__ void __launch_bounds__(kAggBlock)
k_seg_find(uint32_t* __restrict__ parent, uint32_t n, uint32_t* __restrict__ cnt, uint2* __restrict__ nd,
           uint32_t* __restrict__ st, uint32_t round, const RoundSeg* __restrict__ segs_old,
           const uint32_t* __restrict__ prev_mod) {
  const uint32_t live = ld_acq(prev_mod);
  uint32_t gfirst = st[kStGiantFirst];
  const uint64_t best = *reinterpret_cast<const uint64_t*>(st + kStBest);
  const uint32_t nseg = st[kStSeg + 2 * ((round - 1) & 1) + 1];
  const uint32_t i = blockIdx.x * kAggBlock + threadIdx.x;
  const bool in = i < nseg;
  RoundSeg sg{0, 0, 0, 0};
  if (in) sg = segs_old[i];
  // block 0 writes the round state even with no segments to walk
  if (!live || (blockIdx.x > 0 && blockIdx.x * kAggBlock >= nseg)) return;
  __shared__ uint32_t t_key[kAggTable], t_cnt[kAggTable];
  for (uint32_t k = threadIdx.x; k < kAggTable; k += kAggBlock) {
    t_key[k] = kEmptyKey;
    t_cnt[k] = 0;
  }
  if (gfirst == kNoRoot && best) gfirst = (uint32_t)best;  // a candidate recorded by the previous round
  const bool act = gfirst != kNoRoot;
  const uint32_t gcur = act ? dsu_find_stable(parent, gfirst) : kNoRoot;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[kStSeg + 2 * (round & 1)] = 0;
    st[kStSeg + 2 * (round & 1) + 1] = 0;
    st[kStGiantRound] = gcur;
    st[kStGiantFirst] = gfirst;
    st[kStActMode] = act ? 1u : 0u;
  }
  __shared__ uint32_t t_base[kAggTable];
  const bool ok = in && sg.root < n;
  uint32_t r = 0;
  if (ok) r = dsu_find_stable(parent, sg.root);
  __syncthreads();
  const bool counted = ok && r != gcur;
  uint32_t my_slot = 0, my_pos = 0;
  if (ok) parent[sg.root] = r;  // path compression of the old root (members are compressed in C)
  if (counted) {
    my_slot = lds_table_slot(t_key, r);
    my_pos = atomicAdd(t_cnt + my_slot, sg.len);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kAggTable; k += kAggBlock)
    if (t_key[k] != kEmptyKey) t_base[k] = atomicAdd(cnt + t_key[k], t_cnt[k]);
  __syncthreads();
  // the run's offset in its new root's segment, minus its old start: a member at old position i
  // goes to off[new root] + i + nd.y
  if (ok) nd[sg.root] = make_uint2(r, counted ? t_base[my_slot] + my_pos - sg.start : 0u);
}

// B: a segment for every new root (an old root that is still a root) outside the giant; cnt is
// cleared here, its last reader.  Giant election and candidate as in k_bv_alloc.
__global__ void __launch_bounds__(kAggBlock)
k_seg_alloc(uint32_t* __restrict__ cnt, uint32_t n, const uint2* __restrict__ nd, uint32_t* __restrict__ off,
            uint32_t* __restrict__ cursor, RoundSeg* __restrict__ segs_new, uint32_t* __restrict__ st, uint32_t round,
            const RoundSeg* __restrict__ segs_old, uint32_t giant_ok, const uint32_t* __restrict__ prev_mod) {
  const uint32_t live = ld_acq(prev_mod);
  __shared__ uint64_t wsum[kAggBlock / kWave];
  __shared__ uint32_t s_base, s_sbase;
  const bool act = st[kStActMode] != 0;
  const uint32_t nseg = st[kStSeg + 2 * ((round - 1) & 1) + 1];
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint32_t i = blockIdx.x * kAggBlock + threadIdx.x;
  const bool in = i < nseg;
  const uint32_t v = in ? segs_old[i].root : 0;
  const bool ok = in && v < n;
  const bool is_root = ok && nd[v].x == v;
  uint32_t c = is_root ? cnt[v] : 0;  // zero for the giant's root: its members are not counted
  if (!live || blockIdx.x * kAggBlock >= nseg) return;  // whole blocks past the segments
  if (c) cnt[v] = 0;
  if (!act && giant_ok && 2ull * c > n) {
    st[kStGiantFirst] = v;
    st[kStGiantRound] = v;
    c = 0;
  } else if (!act && giant_ok && c >= 2 && c >= n / kCandidateDiv) {
    atomicMax(reinterpret_cast<unsigned long long*>(st + kStBest), ((unsigned long long)c << 32) | v);
  }
  const uint64_t x = c ? ((uint64_t)1 << kSegShift) | c : 0;
  const uint64_t incl = wave_incl_scan_u64(x);
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
  for (uint32_t w = 0; w < kAggBlock / kWave; ++w) {
    if (w < wid) pre += wsum[w];
    tot += wsum[w];
  }
  constexpr uint64_t kCntMask = ((uint64_t)1 << kSegShift) - 1;
  if (threadIdx.x == 0) {
    // entries and supernodes of the block in one 64-bit atomic (neither sum can reach 2^32)
    const uint64_t add = ((tot >> kSegShift) << 32) | (tot & kCntMask);
    const uint64_t old =
        add ? atomicAdd(reinterpret_cast<unsigned long long*>(st + kStSeg + 2 * (round & 1)), (unsigned long long)add) : 0;
    s_base = (uint32_t)old;
    s_sbase = (uint32_t)(old >> 32);
  }
  __syncthreads();
  if (c) {
    const uint64_t ex = pre + incl - x;
    const uint32_t o = s_base + (uint32_t)(ex & kCntMask);
    const uint32_t k = s_sbase + (uint32_t)(ex >> kSegShift);
    off[v] = o;
    cursor[v] = o + c;  // the segment end, as the vertex scatter leaves it
    if (k < n) segs_new[k] = RoundSeg{v, o, c, 0};
  }
}

// C: every member of the old list to its place, its new root beside it, its parent compressed to
// that root.
__global__ void __launch_bounds__(256)
k_seg_copy(uint32_t n, const uint2* __restrict__ nd, const uint32_t* __restrict__ off, uint32_t* __restrict__ parent,
           const uint32_t* __restrict__ members_old, const uint32_t* __restrict__ mroot_old,
           uint32_t* __restrict__ members_new, uint32_t* __restrict__ mroot_new, const uint32_t* __restrict__ st,
           uint32_t round, const uint32_t* __restrict__ prev_mod) {
  const uint32_t live = ld_acq(prev_mod);
  const uint32_t gr = st[kStGiantRound];
  const uint32_t e_old = st[kStSeg + 2 * ((round - 1) & 1)];
  const uint32_t e_new = st[kStSeg + 2 * (round & 1)];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < e_old && i < n;
  uint32_t v = 0, R = 0;
  if (in) {
    v = members_old[i];
    R = mroot_old[i];
  }
  if (!live || !in || R >= n || v >= n) return;
  const uint2 t = nd[R];
  if (t.x == gr || t.x >= n) return;
  const uint32_t p = off[t.x] + i + t.y;
  if (p >= e_new || p >= n) return;
  members_new[p] = v;
  mroot_new[p] = t.x;
  parent[v] = t.x;
}

// Segmented XOR-reduce + sample (range_merge of column `round` + the deterministic bucket over
// every member of a supernode, then sample_supernode).  Entries are the member list, grouped by
// root.  A wave takes 64 consecutive entries; lane l < D accumulates depth l of the column and lane
// D the deterministic bucket.  A supernode whose entries all lie inside the range is sampled from
// registers; its sample record goes to samp[root] for k_apply_samples.  A supernode that crosses
// the range's first (head) or last (tail) boundary leaves its partial XOR in the range's head or
// tail slot under its root; a slot left unused gets root kNoRoot and a zero value.
//
// A workgroup of 16 waves covers 1024 entries: the waves leave their slots in LDS, then one wave
// reduces those 32 slots the same way, so at most two supernodes of a workgroup are unfinished:
// the one continuing from the previous workgroup and the one continuing into the next.
//
// Cross-workgroup supernodes.  A supernode whose entries span workgroups f..l (f < l) is combined
// with device-scope XOR atomics into the accumulator slot of f, its first workgroup (only one
// supernode per workgroup can continue past its end, so f identifies it).  Every workgroup of f..l
// adds its part once and then counts itself in slot f's counter; the one that brings the count to
// l - f + 1 samples the sum and clears the slot and the counter for the next round.  The segment
// bounds come from k_bv_alloc / k_bv_scatter (off = start, cursor = end).  No workgroup waits for
// another, and the XOR is order-free, so the sum is exactly the member-wise merge.
//
// Giant supernode.  Every edge update XORs the same buckets into both endpoints' sketches, so the
// XOR of all vertex sketches is zero and a supernode's sum equals the XOR of every vertex outside
// it.  When a supernode is summed from its complement (the sketches known to sum to zero) its
// members are not regrouped: each workgroup XORs all of its entries into a total and adds it to
// the giant's accumulator slot (slot gridDim.x), counted like the others, and the last workgroup
// samples the giant from that sum -- the same bytes the member-wise merge would give.
constexpr uint32_t kRedWaves = 16, kRedBlock = kRedWaves * kWave;

// Slot of a range's boundary partial: 128 u64 (alpha of lanes 0..63, then gamma) and a root.
struct Slot {
  uint64_t* vals;
  uint32_t* root;
};

// Cross-workgroup accumulators: nslots slots of 128 u64 and as many arrival counters, all zero
// between rounds.  Slots 0 .. gridDim.x - 1 belong to the workgroups, slot nslots - 1 to the giant;
// the launcher checks gridDim.x + 1 <= nslots.
struct XAcc {
  uint64_t* vals;
  uint32_t* cnt;
  uint32_t nslots;
  const uint32_t* seg_off;  // segment start per root
  const uint32_t* seg_end;  // segment end per root (the scatter cursor after k_bv_scatter)
};

// Reduce entries [p0, p0 + len) (len <= 64) with one wave.  prev_root / next_root: roots of the
// entries just before and after the range (kNoRoot when none), which decide whether the first and
// last segments are open.  FIRST: entries are members (rows of the slab); otherwise partial
// entries (in_part / in_root in LDS).  FIRST with ROOTPOS: the entries' roots by position in
// root_pos (members and root_pos may be in LDS) instead of by vertex in vroot.  *all_a / *all_g get the XOR
// of every entry read.  FIRST
// with s_bnd: the wave holding the workgroup's first entry [blk_b0] loads that entry's segment
// bounds into s_bnd[0..1], the one holding its last entry the end of that entry's segment into
// s_bnd[2], in flight with the values.
template <class G, bool FIRST, bool ROOTPOS = false, bool SKIP = false>
__device__ inline void reduce_range(uint32_t p0, uint32_t len, uint32_t prev_root, uint32_t next_root,
                                    const uint32_t* __restrict__ members, const uint32_t* __restrict__ vroot,
                                    const unsigned char* __restrict__ slab, const Geo& g, uint32_t round,
                                    const uint64_t* in_part, const uint32_t* in_root, Slot head, Slot tail,
                                    uint64_t* __restrict__ samp, uint32_t* __restrict__ flags, uint64_t* all_a_out,
                                    uint64_t* all_g_out, const unsigned char* __restrict__ det_cpy,
                                    const XAcc* xa, uint32_t blk_b0, uint32_t blk_b1, uint32_t* s_bnd,
                                    const uint32_t* root_pos = nullptr, bool deep_on = true) {
  const uint32_t lane = lane_id();
  uint32_t m = 0, rt = kNoRoot;
  if (lane < len) {
    if (FIRST) {
      m = members[p0 + lane];
      rt = ROOTPOS ? root_pos[p0 + lane] : vroot[m];
    } else {
      rt = in_root[p0 + lane];
    }
  }
  // SKIP with deep_on (members only): each member's deep flag of column `round`; its deep records
  // are read only when it is set.  The flag shares a line with the column's low records, which the
  // batch loads below read anyway.
  uint32_t deep = 1;
  if (SKIP && FIRST && deep_on && lane < len)
    deep = *reinterpret_cast<const uint32_t*>(slab + (size_t)m * g.row + deep_flag_off(round));
  const uint32_t first_root = rd_lane_u32(rt, 0), last_root = rd_lane_u32(rt, len - 1);
  const bool head_open = first_root != kNoRoot && prev_root == first_root;
  const bool tail_open = last_root != kNoRoot && next_root == last_root;
  const uint32_t p1 = p0 + len;
  const bool holds_first = FIRST && s_bnd && p0 == blk_b0;
  const bool holds_last = FIRST && s_bnd && p1 == blk_b1;
  uint32_t bnd = 0;
  if (holds_first && lane == 0) bnd = xa->seg_off[first_root];
  if (holds_first && lane == 1) bnd = xa->seg_end[first_root];
  if (holds_last && lane == 2) bnd = xa->seg_end[last_root];

  const bool in_col = lane < g.D;
  const uint32_t coff = in_col ? rec_col_off<G>(g, round, lane) : 0;  // depth `lane` of column `round`
  const bool loads = lane <= g.D;
  const bool deep_lane = SKIP && deep_on && in_col && rec_is_deep(g, lane);
  bool wrote_head = false, wrote_tail = false;

  // Every lane stores the root (the same value), so every lane of this wave may read it back
  // without a barrier.
  auto put = [&](Slot sl, uint32_t R, uint64_t a, uint64_t gm) {
    sl.vals[lane] = a;
    sl.vals[kWave + lane] = gm;
    *sl.root = R;
  };
  auto emit = [&](uint32_t R, uint64_t acc_a, uint64_t acc_g, uint32_t seg_start, uint32_t seg_end) {
    const bool started_before = seg_start == p0 && head_open;
    const bool continues = seg_end == p1 && tail_open;
    if (!started_before && !continues) {
      if (R != kNoRoot) {
        uint64_t idx;
        const int code = wave_sample<G>(acc_a, (G)acc_g, g.D, g.chk_bf, &idx);
        if (lane == 0) samp[R] = sample_record(code, idx, g.n, flags);
      }
      return;
    }
    // Partial of a spanning supernode: the head slot if it started before this range, else the
    // tail slot.  When it covers the whole range the tail slot carries the same root with a zero
    // value, so the supernode's entries stay contiguous in the next reduce.
    if (started_before) {
      put(head, R, acc_a, acc_g);
      wrote_head = true;
      if (continues) {
        put(tail, R, 0, 0);
        wrote_tail = true;
      }
    } else {
      put(tail, R, acc_a, acc_g);
      wrote_tail = true;
    }
  };

  uint64_t acc_a = 0, acc_g = 0, all_a = 0, all_g = 0;
  uint32_t cur = first_root, seg_start = p0;
  constexpr uint32_t U = FIRST ? 8 : 16;  // loads in flight per lane
  for (uint32_t j = 0; j < len; j += U) {
    uint64_t va[U], vg[U];
#pragma unroll
    for (uint32_t k = 0; k < U; ++k) {
      va[k] = 0;
      vg[k] = 0;
      if (j + k < len && loads && !(deep_lane && rd_lane_u32(deep, j + k) == 0)) {
        if (FIRST) {
          // lanes < D read the column's records, lane D the compact deterministic bucket: one
          // load instruction with per-lane addresses
          const uint32_t mv = rd_lane_u32(m, j + k);
          const unsigned char* src = in_col ? slab + (size_t)mv * g.row + coff
                                            : reinterpret_cast<const unsigned char*>(det_cpy) + (size_t)mv * (8 + sizeof(G));
          G gw;
          ld_rec<G>(src, &va[k], &gw);
          vg[k] = (uint64_t)gw;
        } else {  // empty slots hold zeros, so the values need not wait for the roots
          const uint64_t* e = in_part + (size_t)(p0 + j + k) * 2 * kWave;
          va[k] = e[lane];
          vg[k] = e[kWave + lane];
        }
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < U; ++k) {
      if (j + k < len) {
        const uint32_t R = rd_lane_u32(rt, j + k);
        if (R != cur) {
          emit(cur, acc_a, acc_g, seg_start, p0 + j + k);
          acc_a = 0;
          acc_g = 0;
          cur = R;
          seg_start = p0 + j + k;
        }
        acc_a ^= va[k];
        acc_g ^= vg[k];
        all_a ^= va[k];
        all_g ^= vg[k];
      }
    }
  }
  emit(cur, acc_a, acc_g, seg_start, p1);
  // an unused slot: no root, zero value (the next reduce reads values without looking at roots)
  if (!wrote_head) put(head, kNoRoot, 0, 0);
  if (!wrote_tail) put(tail, kNoRoot, 0, 0);
  if (holds_first && lane < 2) s_bnd[lane] = bnd;
  if (holds_last && lane == 2) s_bnd[2] = bnd;
  *all_a_out = all_a;
  *all_g_out = all_g;
}

// Per lane: XOR (a, gm) into an accumulator slot with device-scope atomics (zero words skipped).
__device__ inline void xacc_add(uint64_t* slot, uint64_t a, uint64_t gm) {
  const uint32_t lane = lane_id();
  if (a) __hip_atomic_fetch_xor(slot + lane, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (gm) __hip_atomic_fetch_xor(slot + kWave + lane, gm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS of one reduce workgroup: 32 slots of 128 u64 and their roots, the giant total, the
// workgroup's two unfinished supernodes and the segment bounds of its first and last entries.
#define GZ_REDUCE_LDS                                             \
  __shared__ uint64_t s_vals[2 * kRedWaves * 2 * kWave];          \
  __shared__ uint32_t s_root[2 * kRedWaves];                      \
  __shared__ uint64_t s_tot[2 * kWave];                           \
  __shared__ uint64_t s_out[2 * 2 * kWave];                       \
  __shared__ uint32_t s_oroot[2];                                 \
  __shared__ uint32_t s_bnd[3];

// The reduce of a whole round: one workgroup per 1024 entries of the member list (e1 entries,
// written by k_bv_scatter), gridDim.x >= ceil(e1 / 1024).
template <class G>
__global__ void __launch_bounds__(kRedBlock)
k_bv_reduce(const uint32_t* __restrict__ st, const uint32_t* __restrict__ members,
            const uint32_t* __restrict__ mroot, const unsigned char* __restrict__ slab, Geo g, uint32_t round,
            XAcc xa, uint64_t* __restrict__ samp, uint32_t* __restrict__ flags,
            const uint32_t* __restrict__ prev_mod, uint32_t* __restrict__ rstats,
            const unsigned char* __restrict__ det_cpy, uint32_t small_walk, uint32_t deep_skip) {
  const uint32_t live = ld_acq(prev_mod);  // checked after the first reads
  GZ_REDUCE_LDS
  const uint32_t giant = st[kStGiantRound];
  const uint32_t e1 = st[kStSeg + 2 * (round & 1)];
  if (!live) return;
  if (rstats && blockIdx.x == 0 && threadIdx.x == 0 && round < kMaxStatRounds) {
    rstats[2 * round] = e1;  // diagnostics: vertices regrouped and the giant's root in this round
    rstats[2 * round + 1] = giant;
  }
  if (giant != kNoRoot && e1 == 0) {  // every vertex is in the giant: its sum is zero
    if (blockIdx.x == 0 && threadIdx.x == 0) samp[giant] = kSampZero;
    return;
  }
  // Entries per wave: 64, or 32 for a round of at most small_walk entries (16 at most
  // small_walk / 8), whose few workgroups would otherwise each wait out a long chain of dependent
  // gathers (the grid holds enough workgroups for each choice).
  const uint32_t W = e1 <= small_walk / 8 ? kWave / 4 : e1 <= small_walk ? kWave / 2 : kWave;
  const uint32_t span = kRedWaves * W;
  const uint32_t blk = blockIdx.x;
  if (blk * span >= e1) return;  // uniform over the workgroup
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint32_t nblk = (e1 + span - 1) / span;
  const uint32_t b0 = blk * span;
  const uint32_t b1 = b0 + span < e1 ? b0 + span : e1;
  auto root_at = [&](uint32_t p) -> uint32_t { return mroot[p]; };
  const bool giant_tot = giant != kNoRoot;
  // neighbours of the workgroup's range, for phase 2; loaded early so they are in flight with phase 1
  const uint32_t blk_prev = (wid == 0 && b0 > 0) ? root_at(b0 - 1) : kNoRoot;
  const uint32_t blk_next = (wid == 0 && b1 < e1) ? root_at(b1) : kNoRoot;
  if (giant_tot)
    for (uint32_t k = threadIdx.x; k < 2 * kWave; k += blockDim.x) s_tot[k] = 0;
  __syncthreads();
  // phase 1: every wave reduces its entries, leaving its boundary partials in LDS slots
  const uint32_t p0 = b0 + wid * W;
  Slot hs{s_vals + (size_t)(2 * wid) * 2 * kWave, s_root + 2 * wid};
  Slot ts{s_vals + (size_t)(2 * wid + 1) * 2 * kWave, s_root + 2 * wid + 1};
  if (p0 < b1) {
    const uint32_t len = b1 - p0 < W ? b1 - p0 : W;
    const uint32_t prev = p0 == 0 ? kNoRoot : root_at(p0 - 1);
    const uint32_t next = p0 + len >= e1 ? kNoRoot : root_at(p0 + len);
    uint64_t wa, wg;
    reduce_range<G, true, true, kDeepFlags<G>>(p0, len, prev, next, members, nullptr, slab, g, round, nullptr,
                                               nullptr, hs, ts, samp, flags, &wa, &wg, det_cpy, &xa, b0, b1, s_bnd,
                                               mroot, deep_skip != 0);
    if (giant_tot) {
      lds_xor(s_tot + lane, wa);
      lds_xor(s_tot + kWave + lane, wg);
    }
  } else {
    hs.vals[lane] = 0;
    hs.vals[kWave + lane] = 0;
    ts.vals[lane] = 0;
    ts.vals[kWave + lane] = 0;
    if (lane == 0) {
      *hs.root = kNoRoot;
      *ts.root = kNoRoot;
    }
  }
  __syncthreads();
  if (wid != 0) return;  // no barrier follows
  // phase 2: wave 0 reduces the workgroup's 32 slots; the supernodes continuing past either end of
  // the range are left in s_out
  Slot oh{s_out, s_oroot}, ot{s_out + 2 * kWave, s_oroot + 1};
  uint64_t wa, wg;
  reduce_range<G, false>(0, 2 * kRedWaves, blk_prev, blk_next, nullptr, nullptr, nullptr, g, round, s_vals, s_root, oh,
                         ot, samp, flags, &wa, &wg, nullptr, nullptr, 0, 0, nullptr);
  if (nblk == 1) {  // one workgroup: nothing continues, and its total is the giant's whole sum
    if (giant_tot) {
      uint64_t idx;
      const int code = wave_sample<G>(s_tot[lane], (G)s_tot[kWave + lane], g.D, g.chk_bf, &idx);
      if (lane == 0) samp[giant] = sample_record(code, idx, g.n, flags);
    }
    return;
  }
  // Cross-workgroup parts: the head supernode started in workgroup fh (its segment start), the
  // tail one starts here; each ends in the workgroup holding its segment's last entry.
  const uint32_t hr = s_oroot[0], tr = s_oroot[1];
  const bool has_h = hr != kNoRoot;
  const bool has_t = tr != kNoRoot && tr != hr;  // a supernode covering the whole range is counted once
  const uint32_t fh = has_h ? s_bnd[0] / span : 0, lh = has_h ? (s_bnd[1] - 1) / span : 0;
  const uint32_t lt = has_t ? (s_bnd[2] - 1) / span : 0;
  const uint32_t gslot = xa.nslots - 1;
  if (has_h) xacc_add(xa.vals + (size_t)fh * 2 * kWave, s_out[lane], s_out[kWave + lane]);
  if (has_t) xacc_add(xa.vals + (size_t)blk * 2 * kWave, s_out[2 * kWave + lane], s_out[3 * kWave + lane]);
  if (giant_tot) xacc_add(xa.vals + (size_t)gslot * 2 * kWave, s_tot[lane], s_tot[kWave + lane]);
  __threadfence();  // the parts are visible before the arrival is counted
  bool last = false;
  if (lane == 0 && has_h) last = atomicAdd(xa.cnt + fh, 1u) == lh - fh;
  if (lane == 1 && has_t) last = atomicAdd(xa.cnt + blk, 1u) == lt - blk;
  if (lane == 2 && giant_tot) last = atomicAdd(xa.cnt + gslot, 1u) == nblk - 1;
  const uint64_t wins = __ballot(last);
  if (!wins) return;
  __threadfence();  // every part counted before ours is visible
  for (uint32_t k = 0; k < 3; ++k) {
    if (!((wins >> k) & 1)) continue;
    const uint32_t slot = k == 0 ? fh : k == 1 ? blk : gslot;
    const uint32_t R = k == 0 ? hr : k == 1 ? tr : giant;
    uint64_t* sv = xa.vals + (size_t)slot * 2 * kWave;
    const uint64_t a = ld_acq(sv + lane), gm = ld_acq(sv + kWave + lane);
    uint64_t idx;
    const int code = wave_sample<G>(a, (G)gm, g.D, g.chk_bf, &idx);
    if (lane == 0) samp[R] = sample_record(code, idx, g.n, flags);
    // cleared for the next round: every other part of this slot has been counted
    st_rlx(sv + lane, (uint64_t)0);
    st_rlx(sv + kWave + lane, (uint64_t)0);
    if (lane == 0) st_rlx(xa.cnt + slot, 0u);
  }
}

// ---- the last rounds in one workgroup ----------------------------------------------------------
// Once a giant supernode exists, a round walks only the vertices outside it, and after a few rounds
// they are few.  k_bv_finish then runs every remaining round inside one workgroup -- the same
// regrouping, reduce, sample and merge as k_bv_find_count .. k_apply_samples, with barriers between
// the steps instead of kernel boundaries -- so a round costs a few round trips instead of five
// dependent launches.  It takes over at round r0 only when round r0 - 1 changed something, a giant
// exists (or was recorded as the next round's candidate) and the walk holds at most kFinishMax
// vertices; otherwise it returns without touching any state and the rounds continue as launches.
// It then sets mod[r] for every modified round, as the launched rounds do, and records the last
// round it ran in mod[nmod - 1] (a word no round uses), so one copy of mod tells the host both.
//
// The workgroup reduces its grouped list in chunks of 64 entries (one wave each, reduce_range).  A
// supernode continuing past a chunk is accumulated in LDS under the chunk where it starts (only one
// supernode per chunk can continue past its end) and sampled after a barrier.  DSU and sample
// loads that an earlier round of this kernel may have changed are device-coherent loads.
constexpr uint32_t kFinishMax = 2048;
constexpr uint32_t kFinishTable = 4096;  // LDS hash table of roots, >= 2 * kFinishMax, power of two
constexpr uint32_t kFinishChunks = kFinishMax / kWave;
constexpr uint32_t kFinishPer = kFinishMax / kRedBlock;  // walked vertices per thread

__device__ inline uint32_t dsu_find_coherent(uint32_t* parent, uint32_t u) {
  uint32_t p = ld_acq(parent + u);
  while (true) {
    const uint32_t gp = ld_acq(parent + p);
    if (gp == p) return p;
    u = p;
    p = gp;
  }
}

__device__ inline uint32_t finish_slot(uint32_t* keys, uint32_t key) {
  uint32_t h = (key * 2654435761u) & (kFinishTable - 1);
  while (true) {
    const uint32_t prev = atomicCAS(keys + h, kEmptyKey, key);
    if (prev == kEmptyKey || prev == key) return h;
    h = (h + 1) & (kFinishTable - 1);
  }
}

template <class G>
__global__ void __launch_bounds__(kRedBlock)
k_bv_finish(DsuDev d, uint32_t* __restrict__ st, const uint32_t* __restrict__ act0,
            const unsigned char* __restrict__ slab, Geo g, uint32_t r0, uint32_t last_round,
            uint64_t* __restrict__ samp, uint32_t* __restrict__ rstats, const unsigned char* __restrict__ det_cpy,
            const unsigned char* __restrict__ hint, uint32_t* __restrict__ rep) {
  __shared__ uint32_t s_key[kFinishTable], s_cnt[kFinishTable];
  __shared__ uint32_t s_mem[kFinishMax], s_mroot[kFinishMax], s_seg[kFinishMax];
  __shared__ uint64_t s_acc[kFinishChunks * 2 * kWave];
  __shared__ uint32_t s_accroot[kFinishChunks];
  __shared__ uint64_t s_slot[kRedWaves * 2 * 2 * kWave];  // per wave: head and tail of its chunk
  __shared__ uint32_t s_slotroot[kRedWaves * 2];
  __shared__ uint64_t s_tot[2 * kWave];
  __shared__ uint32_t s_wsum[kRedWaves];
  __shared__ uint32_t s_take, s_gfirst, s_walk, s_gcur, s_e;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  if (tid == 0) {
    const uint32_t live = ld_acq(d.mod + (r0 - 1));
    uint32_t gfirst = st[kStGiantFirst];
    const uint64_t best = *reinterpret_cast<const uint64_t*>(st + kStBest);
    if (gfirst == kNoRoot && best) gfirst = (uint32_t)best;  // as k_bv_find_count takes the candidate
    const uint32_t walk = st[kStSeg + 2 * ((r0 - 1) & 1)];
    s_take = (live && gfirst != kNoRoot && walk <= kFinishMax) ? 1u : 0u;
    s_gfirst = gfirst;
    s_walk = walk;
  }
  // The last kernel of every group of rounds: on exit it copies the round flags and the per-round
  // diagnostics to rep (mapped host memory), where the host reads them after its wait, so no copy
  // launch follows the group.
  auto report = [&]() {
    __syncthreads();
    for (uint32_t t = tid; t < d.nmod; t += kRedBlock) rep[t] = ld_acq(d.mod + t);
    for (uint32_t t = tid; t < 2 * kMaxStatRounds; t += kRedBlock) rep[d.nmod + t] = ld_acq(rstats + t);
    __threadfence_system();
  };
  __syncthreads();
  if (!s_take) {
    if (rep) report();
    return;
  }
  uint32_t walk = s_walk;
  uint32_t v[kFinishPer];
#pragma unroll
  for (uint32_t k = 0; k < kFinishPer; ++k) v[k] = tid + k * kRedBlock < walk ? act0[tid + k * kRedBlock] : 0u;
  for (uint32_t c = tid; c < kFinishChunks * 2 * kWave; c += kRedBlock) s_acc[c] = 0;
  for (uint32_t c = tid; c < kFinishChunks; c += kRedBlock) s_accroot[c] = kNoRoot;

  for (uint32_t r = r0;; ++r) {
    for (uint32_t k = tid; k < kFinishTable; k += kRedBlock) {
      s_key[k] = kEmptyKey;
      s_cnt[k] = 0;
    }
    if (tid < 2 * kWave) s_tot[tid] = 0;
    if (tid == 0) s_gcur = dsu_find_coherent(d.parent, s_gfirst);  // the giant only grows
    __syncthreads();
    const uint32_t gcur = s_gcur;
    // roots of the walked vertices, full path compression, members per root (k_bv_find_count)
    uint32_t rr[kFinishPer], slot[kFinishPer], pos[kFinishPer];
#pragma unroll
    for (uint32_t k = 0; k < kFinishPer; ++k) rr[k] = tid + k * kRedBlock < walk ? dsu_find_coherent(d.parent, v[k]) : kNoRoot;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kFinishPer; ++k) {
      if (rr[k] == kNoRoot) continue;
      st_rlx(d.parent + v[k], rr[k]);
      if (rr[k] != gcur) {  // members of the giant are not regrouped
        slot[k] = finish_slot(s_key, rr[k]);
        pos[k] = atomicAdd(s_cnt + slot[k], 1u);
      }
    }
    __syncthreads();
    // segment starts: exclusive scan of the counts over the table (4 slots per thread)
    {
      constexpr uint32_t per = kFinishTable / kRedBlock;
      uint32_t c[per], sum = 0;
#pragma unroll
      for (uint32_t k = 0; k < per; ++k) {
        c[k] = s_cnt[tid * per + k];
        sum += c[k];
      }
      const uint32_t incl = (uint32_t)wave_incl_scan_u64(sum);
      if (lane == kWave - 1) s_wsum[wid] = incl;
      __syncthreads();
      uint32_t pre = 0, tot = 0;
      for (uint32_t w = 0; w < kRedWaves; ++w) {
        if (w < wid) pre += s_wsum[w];
        tot += s_wsum[w];
      }
      uint32_t o = pre + incl - sum;
#pragma unroll
      for (uint32_t k = 0; k < per; ++k) {
        s_cnt[tid * per + k] = o;  // now the segment start
        o += c[k];
      }
      if (tid == 0) s_e = tot;
    }
    __syncthreads();
    const uint32_t e = s_e;
#pragma unroll
    for (uint32_t k = 0; k < kFinishPer; ++k) {
      if (rr[k] == kNoRoot || rr[k] == gcur) continue;
      const uint32_t p = s_cnt[slot[k]] + pos[k];
      s_mem[p] = v[k];
      s_mroot[p] = rr[k];
      s_seg[p] = s_cnt[slot[k]];
    }
    if (tid == 0 && rstats && r < kMaxStatRounds) {
      rstats[2 * r] = e;
      rstats[2 * r + 1] = gcur;
    }
    __syncthreads();
    // reduce + sample (k_bv_reduce): one wave per chunk of cw <= 64 entries, the entries spread
    // over all waves (at least 8 per chunk), so a small round is a short chain of gathers; at most
    // 16 chunks below 64 entries each, at most kFinishChunks of 64
    const uint32_t per_wave = (e + kRedWaves - 1) / kRedWaves;
    const uint32_t cw = per_wave >= kWave ? kWave : (per_wave < 8 ? 8u : (per_wave + 7) & ~7u);
    const uint32_t nchunks = (e + cw - 1) / cw;
    Slot hs{s_slot + (size_t)(2 * wid) * 2 * kWave, s_slotroot + 2 * wid};
    Slot ts{s_slot + (size_t)(2 * wid + 1) * 2 * kWave, s_slotroot + 2 * wid + 1};
    for (uint32_t c = wid; c < nchunks; c += kRedWaves) {
      const uint32_t p0 = c * cw;
      const uint32_t len = e - p0 < cw ? e - p0 : cw;
      const uint32_t prev = p0 > 0 ? s_mroot[p0 - 1] : kNoRoot;
      const uint32_t next = p0 + len < e ? s_mroot[p0 + len] : kNoRoot;
      uint64_t wa, wg;
      reduce_range<G, true, true>(p0, len, prev, next, s_mem, nullptr, slab, g, r, nullptr, nullptr, hs, ts, samp, d.flags,
                            &wa, &wg, det_cpy, nullptr, 0, 0, nullptr, s_mroot);
      lds_xor(s_tot + lane, wa);
      lds_xor(s_tot + kWave + lane, wg);
      const uint32_t hr = *hs.root, tr = *ts.root;
      if (hr != kNoRoot) {  // started in an earlier chunk
        uint64_t* acc = s_acc + (size_t)(s_seg[p0] / cw) * 2 * kWave;
        lds_xor(acc + lane, hs.vals[lane]);
        lds_xor(acc + kWave + lane, hs.vals[kWave + lane]);
      }
      if (tr != kNoRoot && tr != hr) {  // starts here and continues
        uint64_t* acc = s_acc + (size_t)c * 2 * kWave;
        lds_xor(acc + lane, ts.vals[lane]);
        lds_xor(acc + kWave + lane, ts.vals[kWave + lane]);
        if (lane == 0) s_accroot[c] = tr;
      }
    }
    __syncthreads();
    for (uint32_t c = wid; c < nchunks; c += kRedWaves) {
      const uint32_t R = s_accroot[c];
      if (R == kNoRoot) continue;
      uint64_t* acc = s_acc + (size_t)c * 2 * kWave;
      uint64_t idx;
      const int code = wave_sample<G>(acc[lane], (G)acc[kWave + lane], g.D, g.chk_bf, &idx);
      if (lane == 0) samp[R] = sample_record(code, idx, g.n, d.flags);
      acc[lane] = 0;  // zero for the next round
      acc[kWave + lane] = 0;
      if (lane == 0) s_accroot[c] = kNoRoot;
    }
    if (wid == kRedWaves - 1) {  // the giant, from the complement
      uint64_t idx;
      const int code = wave_sample<G>(s_tot[lane], (G)s_tot[kWave + lane], g.D, g.chk_bf, &idx);
      if (lane == 0) samp[gcur] = e == 0 ? kSampZero : sample_record(code, idx, g.n, d.flags);
    }
    // The samples were written with plain stores; the merges read them with device-coherent loads,
    // which a workgroup barrier alone does not order after those stores.
    __threadfence();
    __syncthreads();
    // merges (k_apply_samples): every supernode root of the walk, and the giant
    bool modified = false;
    for (uint32_t k = tid; k < kFinishTable; k += kRedBlock)
      if (s_key[k] != kEmptyKey && apply_sample(d, ld_acq(samp + s_key[k]), hint)) modified = true;
    if (tid == 0 && apply_sample(d, ld_acq(samp + gcur), hint)) modified = true;
    const bool any = __syncthreads_or(modified);
    if (tid == 0 && any) st_rlx(d.mod + r, 1u);
    if (!any || r >= last_round) {
      if (tid == 0) st_rlx(d.mod + (d.nmod - 1), r);
      if (rep) report();
      return;
    }
    // the next round walks this round's members
    walk = e;
#pragma unroll
    for (uint32_t k = 0; k < kFinishPer; ++k) v[k] = tid + k * kRedBlock < walk ? s_mem[tid + k * kRedBlock] : 0u;
    __syncthreads();
  }
}

// ConnectedComponents (src/return_types.cpp:5-19): root of every vertex and the root count.
__global__ void k_labels(uint32_t* __restrict__ parent, uint32_t n, uint32_t* __restrict__ labels,
                         uint32_t* __restrict__ ncc) {
  const uint32_t lane = lane_id();
  for (uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) - lane; v0 < n; v0 += gridDim.x * blockDim.x) {
    const uint32_t v = v0 + lane;
    bool is_root = false;
    if (v < n) {
      const uint32_t r = dsu_find(parent, v);
      labels[v] = r;
      is_root = r == v;
    }
    const uint64_t m = __ballot(is_root);
    if (lane == 0 && m) atomicAdd(ncc, (uint32_t)__popcll(m));
  }
}
__global__ void k_point_queries(uint32_t* __restrict__ parent, const uint32_t* __restrict__ qa,
                                const uint32_t* __restrict__ qb, uint32_t nq, uint8_t* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    out[i] = dsu_find(parent, qa[i]) == dsu_find(parent, qb[i]) ? 1 : 0;
}

// The point queries of a query group and the error flags, read from and written to mapped host
// memory directly (a few words, so no copy launches around the kernel).
__global__ void k_answer_points(uint32_t* __restrict__ parent, const uint32_t* __restrict__ flags,
                                const uint32_t* __restrict__ q, uint32_t nq, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kNumFlags) out[t] = flags[t];
  uint8_t* ans = reinterpret_cast<uint8_t*>(out + kNumFlags);
  for (uint32_t i = t; i < nq; i += gridDim.x * blockDim.x)
    ans[i] = dsu_find(parent, q[i]) == dsu_find(parent, q[nq + i]) ? 1 : 0;
  __threadfence_system();
}

// ================================================================================================
// Standalone sketch ops
// ================================================================================================
// Sketch::sample for any cols_per_sample: deterministic bucket first, then the sample's columns
// in bucket order; one wave walks them 64 buckets at a time.
template <class G>
__global__ void k_sketch_sample(const unsigned char* __restrict__ slab, Geo g, uint32_t sample_idx,
                                uint32_t* __restrict__ out) {
  const uint32_t lane = lane_id();
  uint64_t det_a;
  G det_g;
  ld_rec<G>(slab + rec_off<G>(g, 0, g.nb - 1), &det_a, &det_g);
  int code;
  uint64_t idx = 0;
  if (det_a == 0 && det_g == 0) {
    code = kZero;
  } else if ((G)xxh3_mix(xxh3_prep(det_a), g.chk_bf) == det_g) {
    code = kGood;
    idx = det_a;
  } else {
    code = kFail;
    const uint32_t b0 = sample_idx * g.cps * g.D, b1 = b0 + g.cps * g.D;
    for (uint32_t b = b0; b < b1; b += kWave) {
      const bool valid = b + lane < b1;
      uint64_t a = 0;
      G gm = 0;
      if (valid) ld_rec<G>(slab + rec_off<G>(g, 0, b + lane), &a, &gm);
      const bool good = valid && (G)xxh3_mix(xxh3_prep(a), g.chk_bf) == gm;
      const uint64_t mask = __ballot(good);
      if (mask) {
        idx = rd_lane_u64(a, __builtin_ctzll(mask));
        code = kGood;
        break;
      }
    }
  }
  if (lane == 0) {
    out[0] = (uint32_t)code;
    out[1] = (uint32_t)idx;
    out[2] = (uint32_t)(idx >> 32);
  }
}

template <class G>
__global__ void k_xor_range(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src, Geo g,
                            uint32_t nsk, uint32_t b0, uint32_t cnt, int with_det) {
  const uint32_t per = cnt + (with_det ? 1 : 0);
  const size_t total = (size_t)nsk * per;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)(i / per), k = (uint32_t)(i % per);
    const uint32_t b = k < cnt ? b0 + k : g.nb - 1;
    const size_t o = rec_off<G>(g, s, b);
    uint64_t da, sa;
    G dg, sg;
    ld_rec<G>(dst + o, &da, &dg);
    ld_rec<G>(src + o, &sa, &sg);
    st_rec<G>(dst + o, da ^ sa, (G)(dg ^ sg));
  }
}

// Packed reference image: record i of sketch v at byte (v * nb + i) * (8 + sizeof(G)); the slab row
// holds the same records, padding and the deep flags (set here for every deep record written
// nonzero; a stale set flag only costs the reduce a read).
template <class G, int MODE>  // 0 = xor image into slab, 1 = overwrite slab, 2 = slab to image
__global__ void k_packed(unsigned char* __restrict__ slab, Geo g, uint32_t v0, uint32_t nv,
                         unsigned char* __restrict__ packed) {
  const size_t rec = 8 + sizeof(G);
  const size_t total = (size_t)nv * g.nb;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(i / g.nb), b = (uint32_t)(i % g.nb);
    unsigned char* r = slab + rec_off<G>(g, v0 + v, b);
    uint64_t a, pa;
    G gm, pg;
    ld_rec<G>(r, &a, &gm);
    if (MODE == 2) {
      st_rec<G>(packed + i * rec, a, gm);
    } else {
      ld_rec<G>(packed + i * rec, &pa, &pg);
      const uint64_t na = MODE == 0 ? a ^ pa : pa;
      const G ng = MODE == 0 ? (G)(gm ^ pg) : pg;
      st_rec<G>(r, na, ng);
      if (kDeepFlags<G> && b != g.nb - 1 && (na | (uint64_t)ng) != 0) {
        const uint32_t c = b / g.D;
        if (rec_is_deep(g, b - c * g.D))
          st_rlx(reinterpret_cast<uint32_t*>(slab + (size_t)(v0 + v) * g.row + deep_flag_off(c)), 1u);
      }
    }
  }
}

// Bucket-wise XOR of every vertex sketch into out_a / out_g (nb entries each, zero on entry): one
// thread per (bucket, chunk of kSumChunk vertices), one device-scope atomic per thread.
constexpr uint32_t kSumChunk = 1024;
template <class G>
__global__ void __launch_bounds__(256)
k_sum_all(const unsigned char* __restrict__ slab, Geo g, uint64_t* __restrict__ out_a, G* __restrict__ out_g) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nb) return;
  const uint32_t v0 = blockIdx.y * kSumChunk;
  const uint32_t v1 = v0 + kSumChunk < g.n ? v0 + kSumChunk : g.n;
  uint64_t a = 0;
  G gm = 0;
  for (uint32_t v = v0; v < v1; ++v) {
    uint64_t ra;
    G rg;
    ld_rec<G>(slab + rec_off<G>(g, v, b), &ra, &rg);
    a ^= ra;
    gm ^= rg;
  }
  if (a) gbl_xor(out_a + b, a);
  if (gm) gbl_xor(out_g + b, gm);
}

// Order-sensitive digest of every vertex's bucket array:
//   D = sum over vertices v, buckets b of fmix64(x + pos), pos = v*nb + b,
//   x = alpha ^ fmix64(gamma ^ pos * 0x9E3779B97F4A7C15)   (all mod 2^64)
// A sum is independent of reduction order, so the value is deterministic.
template <class G>
__global__ void __launch_bounds__(256)
k_digest(const unsigned char* __restrict__ slab, Geo g, unsigned long long* __restrict__ out) {
  const size_t total = (size_t)g.n * g.nb;
  uint64_t acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(i / g.nb), b = (uint32_t)(i % g.nb);
    uint64_t a;
    G gm;
    ld_rec<G>(slab + rec_off<G>(g, v, b), &a, &gm);
    const uint64_t pos = i;
    const uint64_t x = a ^ mix64((uint64_t)gm ^ (pos * 0x9E3779B97F4A7C15ULL));
    acc += mix64(x + pos);
  }
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((unsigned)(uint32_t)acc, o), hi = __shfl_xor((unsigned)(uint32_t)(acc >> 32), o);
    acc += ((uint64_t)hi << 32) | lo;
  }
  if (lane_id() == 0) atomicAdd(out, (unsigned long long)acc);
}

}  // namespace

// ================================================================================================
// launchers
// ================================================================================================
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}

size_t update_tiles_lds_bytes(const Geo& g, int gbytes) {
  const size_t lds_n = (size_t)g.C * kLdsPitch + 1;
  return kWave * 16 + g.C * 8 + lds_n * 8 + lds_n * (size_t)gbytes + lds_n * 4;
}

void launch_update_tiles(int gbytes, const Tile* tiles, const uint32_t* ntiles_dev, uint32_t ntiles_host,
                         uint32_t grid_cap, const uint32_t* dsts, const uint64_t* raw_idx, unsigned char* slab,
                         const Geo& g, const uint64_t* col_bf, hipStream_t s) {
  // tile count known on the host: one workgroup per tile (capped, then grid-stride);
  // known only on the device: a persistent grid of grid_cap workgroups
  uint32_t grid = ntiles_dev ? grid_cap : (grid_cap && ntiles_host > grid_cap ? grid_cap : ntiles_host);
  if (grid == 0) return;
  const size_t lds = update_tiles_lds_bytes(g, gbytes);
  const bool wide = g.C * kLdsPitch + 1 > 4 * kWave;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWave), lds, s, tiles, ntiles_dev, ntiles_host, dsts, raw_idx, slab, g,
                       col_bf);
  };
  if (gbytes == 4) {
    if (raw_idx) wide ? go(k_update_tiles<uint32_t, true, true>) : go(k_update_tiles<uint32_t, true, false>);
    else wide ? go(k_update_tiles<uint32_t, false, true>) : go(k_update_tiles<uint32_t, false, false>);
  } else {
    if (raw_idx) wide ? go(k_update_tiles<uint64_t, true, true>) : go(k_update_tiles<uint64_t, true, false>);
    else wide ? go(k_update_tiles<uint64_t, false, true>) : go(k_update_tiles<uint64_t, false, false>);
  }
}

size_t group_partials_len(uint32_t n) { return (n + kScanSpan - 1) / kScanSpan + 1; }

void launch_group(const uint32_t* pairs, size_t count, uint32_t n, uint32_t tile, const GroupBufs& b, hipStream_t s,
                  const FlushBytes* fb) {
  const uint2* p2 = reinterpret_cast<const uint2*>(pairs);
  uint32_t shift = 5;
  while ((size_t(1) << shift) * kPartBins < n) ++shift;
  if (b.rec && (1u << shift) <= kPartMaxRange) {
    const uint32_t bins = (uint32_t)((n + (1u << shift) - 1) >> shift);
    const uint32_t nblk = (uint32_t)std::max<size_t>(1, std::min<size_t>(kPartMaxBlocks, (count + 4095) / 4096));
    const uint32_t hn = bins * nblk;
    if (count) {
      hipLaunchKernelGGL(k_part_hist, dim3(nblk), dim3(kPartBlock), 0, s, p2, count, n, shift, bins, nblk, b.hist,
                         b.bad);
    } else {
      launch_fill_u32(b.hist, 0, hn, s);
    }
    scan_counts(b.hist, hn, 0, b.hpartials, b.hoff, nullptr, nullptr, b.htotals, s);
    if (count)
      hipLaunchKernelGGL(k_part_scatter, dim3(nblk), dim3(kPartBlock), 0, s, p2, count, n, shift, bins, nblk, b.hoff,
                         b.rec);
    hipLaunchKernelGGL(k_part_group, dim3(bins), dim3(kPartBlock), 0, s, b.rec, n, shift, bins, nblk, b.hoff, b.cnt,
                       b.dst_arr);
    if (scan_counts(b.cnt, n, tile, b.partials, b.off, b.toff, nullptr, b.totals, s, b.tiles, fb)) return;
  } else {
    launch_fill_u32(b.cnt, 0, n, s);
    if (count) hipLaunchKernelGGL(k_count_pairs, dim3(grid_for(count, 256)), dim3(256), 0, s, p2, count, n, b.cnt, b.bad);
    scan_counts(b.cnt, n, tile, b.partials, b.off, b.toff, b.cursor, b.totals, s);
    if (count) hipLaunchKernelGGL(k_scatter_pairs, dim3(grid_for(count, 256)), dim3(256), 0, s, p2, count, n, b.cursor, b.dst_arr);
  }
  hipLaunchKernelGGL(k_make_tiles, dim3(grid_for(n, 256, kProfSlots)), dim3(256), 0, s, b.cnt, b.off, b.toff, n, tile,
                     b.tiles, fb ? fb->acc : nullptr, fb ? fb->S : 0, fb ? fb->per_update : 0);
}

size_t part_hist_len() { return (size_t)kPartBins * kPartMaxBlocks + 1; }

void launch_eager(const uint32_t* src, uint32_t* pairs, size_t count, uint8_t* merged, const DsuDev& d,
                  hipStream_t s, bool build_sf) {
  if (!count) return;
  uint2* p2 = reinterpret_cast<uint2*>(pairs);
  const uint32_t grid = grid_for(count, 256);
  const uint32_t ugrid = build_sf ? std::max(grid, grid_for(d.n, 256, 4096)) : grid;
  hipLaunchKernelGGL(k_eager_union, dim3(ugrid), dim3(256), 0, s, reinterpret_cast<const uint2*>(src), p2, count,
                     merged, d, build_sf ? 1u : 0u);
  hipLaunchKernelGGL(k_eager_sf_check, dim3(grid), dim3(256), 0, s, p2, count, merged, d);
}


void launch_dsu_reset(const DsuDev& d, hipStream_t s, const RoundBufs* rb) {
  const size_t work = std::max<size_t>(d.n, d.sf_mask + 1);
  hipLaunchKernelGGL(k_dsu_reset, dim3(grid_for(work, 256)), dim3(256), 0, s, d.parent, d.n, d.sf_hash,
                     d.sf_mask + 1, d.sf_edge, d.flags, d.mod, d.nmod, rb ? rb->cnt : nullptr,
                     rb ? rb->totals : nullptr, rb ? d.query_force : nullptr);
}

void launch_round_state_init(const RoundBufs& rb, uint32_t n, hipStream_t s) {
  launch_fill_u32(rb.totals, 0, kStWords, s);
  launch_fill_u32(rb.totals + kStGiantRound, kNoRoot, 1, s);
  launch_fill_u32(rb.totals + kStGiantFirst, kNoRoot, 1, s);
  launch_fill_u32(rb.members[0], 0, n, s);
  launch_fill_u32(rb.members[1], 0, n, s);
  launch_fill_u32(rb.mroot[0], 0, n, s);
  launch_fill_u32(rb.mroot[1], 0, n, s);
}

void launch_sample_round0(int gbytes, const unsigned char* slab, const Geo& g, const DsuDev& d, const RoundBufs& rb,
                          hipStream_t s) {
  uint64_t* samp = rb.samp;
  const uint32_t grid = grid_for((size_t)(g.n + 3) / 4 * kWave, 256, 16384);
  if (gbytes == 4)
    hipLaunchKernelGGL(k_sample_round0<uint32_t>, dim3(grid), dim3(256), 0, s, slab, g, samp, d.flags, rb.det_cpy,
                       rb.hint, d);
  else
    hipLaunchKernelGGL(k_sample_round0<uint64_t>, dim3(grid), dim3(256), 0, s, slab, g, samp, d.flags, rb.det_cpy,
                       rb.hint, d);
  hipLaunchKernelGGL(k_apply_samples, dim3(grid_for(g.n, 256, 4096)), dim3(256), 0, s, samp, nullptr, d, nullptr, 0u,
                     nullptr, d.mod, rb.hint);
}

// Host bound on each level's entry count (the kernels compute the exact counts on the device):
// the larger of the chain without a supernode summed from its complement (level 2 = 2 * B1) and
// with one (level 2 = 3 * B1); e1 <= n in both.
uint32_t reduce_blocks(uint32_t n) { return 2 * std::max(1u, (n + kRedBlock - 1) / kRedBlock); }
constexpr uint32_t kSmallWalkDiv = 2;  // rounds regrouping at most n / kSmallWalkDiv entries use 32 per wave

void launch_round(int gbytes, uint32_t round, const unsigned char* slab, const Geo& g,
                  const DsuDev& d, const RoundBufs& rb, bool giant_ok, hipStream_t s, bool deep_skip) {
  const uint32_t n = g.n;
  const uint32_t* prev = d.mod + (round - 1);
  uint32_t* cur = d.mod + round;
  uint32_t* st = rb.totals;
  const uint32_t par = round & 1, opar = (round - 1) & 1;
  uint32_t* members = rb.members[par];
  const uint32_t grid = (n + kAggBlock - 1) / kAggBlock;  // one vertex, or one old segment, per thread
  if (round == 1) {  // every vertex in id order
    const uint32_t* act_list = rb.members[opar];  // not read in round 1
    hipLaunchKernelGGL(k_bv_find_count, dim3(grid), dim3(kAggBlock), 0, s, d.parent, n, rb.nd, rb.cnt, st, round,
                       act_list, prev);
    hipLaunchKernelGGL(k_bv_alloc, dim3(grid), dim3(kAggBlock), 0, s, rb.cnt, n, rb.off, rb.cursor, st, round,
                       act_list, giant_ok ? 1u : 0u, prev, rb.segs[par]);
    hipLaunchKernelGGL(k_bv_scatter, dim3(grid), dim3(kAggBlock), 0, s, rb.nd, n, rb.off, members, rb.mroot[par],
                       st, round, act_list, prev);
  } else {  // whole runs of the previous round's member list
    hipLaunchKernelGGL(k_seg_find, dim3(grid), dim3(kAggBlock), 0, s, d.parent, n, rb.cnt, rb.nd, st, round,
                       rb.segs[opar], prev);
    hipLaunchKernelGGL(k_seg_alloc, dim3(grid), dim3(kAggBlock), 0, s, rb.cnt, n, rb.nd, rb.off, rb.cursor,
                       rb.segs[par], st, round, rb.segs[opar], giant_ok ? 1u : 0u, prev);
    hipLaunchKernelGGL(k_seg_copy, dim3(grid_for(n, 256)), dim3(256), 0, s, n, rb.nd, rb.off, d.parent, rb.members[opar],
                       rb.mroot[opar], members, rb.mroot[par], st, round, prev);
  }
  // one accumulator slot per workgroup and one for the giant
  const uint32_t rgrid = reduce_blocks(n);
  if (rgrid + 1 > rb.xslots) throw std::logic_error("reduce grid exceeds its cross-workgroup accumulator slots");
  const XAcc xa{rb.xacc, rb.xcnt, rb.xslots, rb.off, rb.cursor};
  if (gbytes == 4)
    hipLaunchKernelGGL(k_bv_reduce<uint32_t>, dim3(rgrid), dim3(kRedBlock), 0, s, st, members, rb.mroot[par],
                       slab, g, round, xa, rb.samp, d.flags, prev, rb.rstats, rb.det_cpy, n / kSmallWalkDiv,
                       deep_skip ? 1u : 0u);
  else
    hipLaunchKernelGGL(k_bv_reduce<uint64_t>, dim3(rgrid), dim3(kRedBlock), 0, s, st, members, rb.mroot[par],
                       slab, g, round, xa, rb.samp, d.flags, prev, rb.rstats, rb.det_cpy, n / kSmallWalkDiv,
                       deep_skip ? 1u : 0u);
  hipLaunchKernelGGL(k_apply_samples, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, rb.samp, rb.segs[par], d, st,
                     round, prev, cur, rb.hint);
}

bool launch_finish(int gbytes, uint32_t r0, uint32_t last_round, const unsigned char* slab, const Geo& g,
                   const DsuDev& d, const RoundBufs& rb, hipStream_t s, uint32_t* rep) {
  if (r0 < 2 || r0 > last_round) return false;  // round r0 walks round r0 - 1's member list
  const uint32_t* act0 = rb.members[(r0 - 1) & 1];
  if (gbytes == 4)
    hipLaunchKernelGGL(k_bv_finish<uint32_t>, dim3(1), dim3(kRedBlock), 0, s, d, rb.totals, act0, slab, g, r0,
                       last_round, rb.samp, rb.rstats, rb.det_cpy, rb.hint, rep);
  else
    hipLaunchKernelGGL(k_bv_finish<uint64_t>, dim3(1), dim3(kRedBlock), 0, s, d, rb.totals, act0, slab, g, r0,
                       last_round, rb.samp, rb.rstats, rb.det_cpy, rb.hint, rep);
  return true;
}
uint32_t finish_max() { return kFinishMax; }


void launch_sum_all(int gbytes, const unsigned char* slab, const Geo& g, uint64_t* out_a, void* out_g, hipStream_t s) {
  launch_fill_u64(out_a, 0, g.nb, s);
  if (gbytes == 4) launch_fill_u32((uint32_t*)out_g, 0, g.nb, s);
  else launch_fill_u64((uint64_t*)out_g, 0, g.nb, s);
  const dim3 grid((g.nb + 255) / 256, (g.n + kSumChunk - 1) / kSumChunk);
  if (gbytes == 4)
    hipLaunchKernelGGL(k_sum_all<uint32_t>, grid, dim3(256), 0, s, slab, g, out_a, (uint32_t*)out_g);
  else
    hipLaunchKernelGGL(k_sum_all<uint64_t>, grid, dim3(256), 0, s, slab, g, out_a, (uint64_t*)out_g);
}

void launch_digest(int gbytes, const unsigned char* slab, const Geo& g, uint64_t* out, hipStream_t s) {
  launch_fill_u64(out, 0, 1, s);
  const uint32_t grid = grid_for((size_t)g.n * g.nb, 256, 4096);
  auto* o = reinterpret_cast<unsigned long long*>(out);
  if (gbytes == 4)
    hipLaunchKernelGGL(k_digest<uint32_t>, dim3(grid), dim3(256), 0, s, slab, g, o);
  else
    hipLaunchKernelGGL(k_digest<uint64_t>, dim3(grid), dim3(256), 0, s, slab, g, o);
}

void launch_labels(const DsuDev& d, uint32_t* labels, uint32_t* ncc, hipStream_t s) {
  launch_fill_u32(ncc, 0, 1, s);
  hipLaunchKernelGGL(k_labels, dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.parent, d.n, labels, ncc);
}
void launch_answer_points(const DsuDev& d, const uint32_t* q, uint32_t nq, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_answer_points, dim3(grid_for(nq > kNumFlags ? nq : kNumFlags, 256)), dim3(256), 0, s, d.parent,
                     d.flags, q, nq, out);
}
void launch_point_queries(const DsuDev& d, const uint32_t* qa, const uint32_t* qb, uint32_t nq, uint8_t* out,
                          hipStream_t s) {
  if (!nq) return;
  hipLaunchKernelGGL(k_point_queries, dim3(grid_for(nq, 256)), dim3(256), 0, s, d.parent, qa, qb, nq, out);
}

void launch_sketch_sample(int gbytes, const unsigned char* slab, const Geo& g, uint32_t sample_idx, uint32_t* out,
                          hipStream_t s) {
  if (gbytes == 4)
    hipLaunchKernelGGL(k_sketch_sample<uint32_t>, dim3(1), dim3(kWave), 0, s, slab, g, sample_idx, out);
  else
    hipLaunchKernelGGL(k_sketch_sample<uint64_t>, dim3(1), dim3(kWave), 0, s, slab, g, sample_idx, out);
}
void launch_xor_range(int gbytes, unsigned char* dst, const unsigned char* src, const Geo& g, uint32_t nsk, uint32_t b0,
                      uint32_t cnt, bool with_det, hipStream_t s) {
  const size_t total = (size_t)nsk * (cnt + (with_det ? 1 : 0));
  if (!total) return;
  if (gbytes == 4)
    hipLaunchKernelGGL(k_xor_range<uint32_t>, dim3(grid_for(total, 256)), dim3(256), 0, s, dst, src, g, nsk, b0, cnt,
                       with_det ? 1 : 0);
  else
    hipLaunchKernelGGL(k_xor_range<uint64_t>, dim3(grid_for(total, 256)), dim3(256), 0, s, dst, src, g, nsk, b0, cnt,
                       with_det ? 1 : 0);
}

template <int MODE>
static void launch_packed(int gbytes, unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, const void* packed,
                          hipStream_t s) {
  const size_t total = (size_t)nv * g.nb;
  if (!total) return;
  unsigned char* p = (unsigned char*)packed;
  if (gbytes == 4)
    hipLaunchKernelGGL((k_packed<uint32_t, MODE>), dim3(grid_for(total, 256)), dim3(256), 0, s, slab, g, v0, nv, p);
  else
    hipLaunchKernelGGL((k_packed<uint64_t, MODE>), dim3(grid_for(total, 256)), dim3(256), 0, s, slab, g, v0, nv, p);
}
void launch_merge_packed(int gbytes, unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, const void* packed,
                         hipStream_t s) {
  launch_packed<0>(gbytes, slab, g, v0, nv, packed, s);
}
void launch_unpack(int gbytes, unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, const void* packed,
                   hipStream_t s) {
  launch_packed<1>(gbytes, slab, g, v0, nv, packed, s);
}
void launch_pack(int gbytes, const unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, void* packed,
                 hipStream_t s) {
  launch_packed<2>(gbytes, const_cast<unsigned char*>(slab), g, v0, nv, packed, s);
}

}  // namespace gz
