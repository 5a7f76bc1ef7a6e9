// gz_kernels.hpp -- host-side launchers for the gfx950 kernels in gz_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "gz_common.hpp"

namespace gz {

// Device-side state of one DSU / spanning-forest instance.
struct DsuDev {
  uint32_t* parent;    // n
  uint64_t* sf_hash;   // open-addressing set of spanning-forest edge ids, cap = sf_mask + 1
  uint64_t sf_mask;
  uint64_t* sf_edge;   // n: forest edge (min << 32 | max) that linked root v under another, or ~0
  uint32_t* flags;     // FlagSlot entries
  uint32_t* mod;       // per-round "modified" flags, mod[r] for round r (nmod entries)
  uint32_t nmod;
  uint32_t prio_seed;
  uint32_t n;
  uint32_t* inv_mirror;  // mapped host word set to 1 with kFlagEagerInvalid (nullptr: none)
  // Mapped host word read at a query's start: nonzero, or a set kFlagEagerInvalid, runs Borůvka;
  // otherwise the query keeps the valid eager DSU (nullptr: always run).
  const uint32_t* query_force;
};

enum FlagSlot {
  kFlagModified = 0,
  kFlagEagerInvalid = 1,
  kFlagSkip = 2,      // the last query found the eager DSU valid and left it in place (no Borůvka)
  kFlagBad = 3,       // a sample decoded to an out-of-range vertex
  kFlagBadInput = 4,  // an ingested pair named a vertex >= n (the pair was skipped)
  kNumFlags = 8
};

// fills (graph-capture safe replacements for hipMemsetAsync)
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s);

// ---- sketch update ----
// Tile kernel (one wave per tile).  `ntiles_dev` (device scalar) bounds the tile index when
// non-null (persistent grid of grid_cap workgroups), otherwise `ntiles_host` (one workgroup per
// tile, at most grid_cap when grid_cap > 0).  Workgroups walk tiles grid-stride.
// Exactly one of `dsts` (destination vertex ids; index = concat(vertex, dst)) or `raw_idx`
// (64-bit indices) is non-null.
void launch_update_tiles(int gbytes, const Tile* tiles, const uint32_t* ntiles_dev, uint32_t ntiles_host,
                         uint32_t grid_cap, const uint32_t* dsts, const uint64_t* raw_idx, unsigned char* slab,
                         const Geo& g, const uint64_t* col_bf, hipStream_t s);
size_t update_tiles_lds_bytes(const Geo& g, int gbytes);

// ---- grouping of a device-resident stream of (src,dst) pairs into per-vertex tiles ----
struct GroupBufs {
  uint32_t* cnt;       // n      per-vertex directed-update counts (zeroed by launch_group)
  uint32_t* off;       // n + 1  exclusive scan of cnt
  uint32_t* toff;      // n + 1  exclusive scan of per-vertex tile counts
  uint32_t* cursor;    // n      scatter cursors
  uint64_t* partials;  // scan block partials
  uint32_t* totals;    // [0] total directed updates, [1] total tiles
  uint32_t* dst_arr;   // 2 * max_pairs
  Tile* tiles;         // max tiles
  uint32_t* bad;       // set to 1 when a pair names a vertex >= n
  // bucket partition (used when non-null and the vertex count allows it)
  uint2* rec;          // 2 * max_pairs records (vertex, other end)
  uint32_t* hist;      // part_hist_len()
  uint32_t* hoff;      // part_hist_len()
  uint64_t* hpartials; // group_partials_len(part_hist_len())
  uint32_t* htotals;   // 4
};
size_t part_hist_len();
size_t group_partials_len(uint32_t n);
// pairs: 2*count u32 (src,dst) in HBM.  Produces tiles of at most `tile` destinations.
// Profiling: per-block accumulators of a grouped flush's algorithmic bytes (kProfSlots u64,
// summed on the host).
constexpr uint32_t kProfSlots = 8192;
struct FlushBytes {
  uint64_t* acc;
  uint64_t S;           // bytes of one vertex sketch
  uint64_t per_update;  // (C + 1) * (8 + sizeof(gamma)) bytes touched by one directed update
};
void launch_group(const uint32_t* pairs, size_t count, uint32_t n, uint32_t tile, const GroupBufs& b, hipStream_t s,
                  const FlushBytes* fb = nullptr);

// ---- eager DSU (batch form of pre_insert, src/cc_sketch_alg.cpp:79-104) ----
// src != pairs: the batch is copied from src into pairs by the same launch.
// build_sf: the same launch first fills the forest membership set from d.sf_edge (the first
// batch after a query).
void launch_eager(const uint32_t* src, uint32_t* pairs, size_t count, uint8_t* merged, const DsuDev& d,
                  hipStream_t s, bool build_sf = false);

// ---- Borůvka ----
struct RoundBufs;
// rb (a query's start): also clears rb->cnt and sets the round state words to their start values.
void launch_dsu_reset(const DsuDev& d, hipStream_t s, const RoundBufs* rb = nullptr);
// One supernode of a round outside the giant: its root and its run [start, start + len) in the
// round's member list.
struct RoundSeg {
  uint32_t root, start, len, pad;
};
struct RoundBufs {
  uint32_t* cnt;         // n      members per root (zero between rounds)
  uint32_t* off;         // n      segment start per root
  uint32_t* cursor;      // n      scatter cursor per root (the segment end after the scatter)
  uint32_t* members[2];  // n each vertices grouped by root, by round parity (also the next active list)
  uint32_t* mroot[2];    // n each the root of members[parity][p] at position p
  RoundSeg* segs[2];     // n each the round's supernodes outside the giant, by round parity
  uint2* nd;             // n      round 1, per vertex: its root and its place in the root's segment;
                         //        rounds >= 2, per previous-round root: its new root, and the shift
                         //        from its old run to its place in the new root's segment
  uint32_t* totals;      // 8      round state words (entry counts, giant roots, walk mode)
  uint64_t* xacc;        // xslots * 128: cross-workgroup supernode sums, zero between rounds
  uint32_t* xcnt;        // xslots arrival counters, zero between rounds
  uint32_t xslots;       // >= reduce_blocks(n) + 1 (one per reduce workgroup, one for the giant)
  uint64_t* samp;        // n      sample record per supernode root
  uint32_t* rstats;      // 2 * 64 per round: vertices regrouped, giant root (diagnostics)
  unsigned char* det_cpy;  // n records: every vertex's deterministic bucket, copied in round 0
  unsigned char* hint;     // n      round 0: each vertex's degree class, the first key of its merge order
};
// Start values of the round state words (no giant, empty walks) and in-range member lists, for a new
// instance: round kernels read both before they look at their round flag.
void launch_round_state_init(const RoundBufs& rb, uint32_t n, hipStream_t s);
// Writes rb.samp: a sample record per vertex.
void launch_sample_round0(int gbytes, const unsigned char* slab, const Geo& g, const DsuDev& d, const RoundBufs& rb,
                          hipStream_t s);
// Workgroups of a round's reduce over n vertices (enough for 512 member entries each).
uint32_t reduce_blocks(uint32_t n);
// Rounds r0, r0 + 1, ... up to last_round in one workgroup, when round r0 - 1 changed something, a
// giant supernode exists and round r0 walks at most finish_max() vertices; otherwise a no-op.
// d.mod[d.nmod - 1] gets the last round it ran (it stays 0 when the finish did not take over).
// rep (mapped host memory, or nullptr): on exit the kernel copies d.mod[0..nmod) and the 128
// per-round diagnostic words there.  Returns whether it launched (2 <= r0 <= last_round).
bool launch_finish(int gbytes, uint32_t r0, uint32_t last_round, const unsigned char* slab, const Geo& g,
                   const DsuDev& d, const RoundBufs& rb, hipStream_t s, uint32_t* rep = nullptr);
uint32_t finish_max();
// Queue round `round` (>= 1); every kernel of it is a no-op when d.mod[round - 1] == 0.
// giant_ok: the vertex sketches are known to XOR to zero, so a supernode holding more than half of
// the vertices may be summed from its complement.  deep_skip: the reduce reads a member's records
// past the first overflow line of the round's column only when its deep flag is set (rows of
// 12-byte records); false reads them all (same result).
void launch_round(int gbytes, uint32_t round, const unsigned char* slab, const Geo& g,
                  const DsuDev& d, const RoundBufs& rb, bool giant_ok, hipStream_t s, bool deep_skip = true);
void launch_labels(const DsuDev& d, uint32_t* labels, uint32_t* ncc, hipStream_t s);
// A query group's answers in place in mapped host memory: q[0..nq) and q[nq..2nq) are the pairs;
// out gets the kNumFlags error flags, then one byte per pair (1: connected).
void launch_answer_points(const DsuDev& d, const uint32_t* q, uint32_t nq, uint32_t* out, hipStream_t s);
void launch_point_queries(const DsuDev& d, const uint32_t* qa, const uint32_t* qb, uint32_t nq, uint8_t* out,
                          hipStream_t s);

// order-sensitive 64-bit digest of all n sketches (see k_digest)
void launch_digest(int gbytes, const unsigned char* slab, const Geo& g, uint64_t* out, hipStream_t s);
// Bucket-wise XOR of all n vertex sketches (out_a: nb u64, out_g: nb gamma words).
void launch_sum_all(int gbytes, const unsigned char* slab, const Geo& g, uint64_t* out_a, void* out_g, hipStream_t s);

// ---- standalone sketch ops ----
// out[0] = code, out[1..2] = idx (lo, hi)
void launch_sketch_sample(int gbytes, const unsigned char* slab, const Geo& g, uint32_t sample_idx, uint32_t* out,
                          hipStream_t s);
// dst ^= src on buckets [b0, b0 + cnt) of `nsk` consecutive sketches, plus the deterministic bucket
void launch_xor_range(int gbytes, unsigned char* dst, const unsigned char* src, const Geo& g, uint32_t nsk, uint32_t b0,
                      uint32_t cnt, bool with_det, hipStream_t s);
// XOR a packed (reference-format) bucket image of `nv` sketches starting at v0 into the slab
void launch_merge_packed(int gbytes, unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, const void* packed,
                         hipStream_t s);
// pack `nv` sketches starting at v0 into the reference's byte image (records {u64 alpha, gamma})
void launch_pack(int gbytes, const unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, void* packed,
                 hipStream_t s);
// unpack (overwrite) a packed image into the slab
void launch_unpack(int gbytes, unsigned char* slab, const Geo& g, uint32_t v0, uint32_t nv, const void* packed,
                   hipStream_t s);

}  // namespace gz
