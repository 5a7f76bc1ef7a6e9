// gz_runtime.cpp -- host runtime behind the C ABI (include/gz_hip.h).
//
// One GzAlg owns one graph stream's state on the device: the sketch slab of every vertex, the DSU
// and spanning-forest set, the Borůvka round buffers and the ingest buffers.  All device work of an
// instance is ordered on its own HIP stream (default priority).
#include "gz_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>

namespace gz {

int device_cu_count() {
  static int cus = -1;
  if (cus < 0) {
    int dev = 0;
    GZ_HIP(hipGetDevice(&dev));
    int v = 0;
    GZ_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
    cus = v > 0 ? v : 1;
  }
  return cus;
}

void require_device() {
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1)
    throw GzError(GZ_ERR_DEVICE, "no HIP device is available: the HIP library needs a GPU");
}

static Geo make_geo(uint32_t nsk, uint64_t vector_len, uint64_t seed, size_t samples, size_t cols, int gb) {
  if (vector_len < 1) throw GzError(GZ_ERR_INVALID_ARG, "vector length must be >= 1");
  if (samples < 1 || cols < 1) throw GzError(GZ_ERR_INVALID_ARG, "num_samples and cols_per_sample must be >= 1");
  Geo g{};
  g.n = nsk;
  const size_t D = calc_bkt_per_col(vector_len);
  if (D < 1 || D > 62) throw GzError(GZ_ERR_INVALID_ARG, "bkt_per_col out of the supported range [1, 62]");
  g.D = (uint32_t)D;
  g.samples = (uint32_t)samples;
  g.cps = (uint32_t)cols;
  g.C = (uint32_t)(samples * cols);
  g.nb = g.C * g.D + 1;
  const RowLayout L = row_layout(g.C, g.D, 8 + gb);
  g.low = L.low;
  g.ovf = L.ovf;
  g.opitch = L.opitch;
  g.row = L.row;
  g.Dp = lds_pitch(g.D);
  g.seed = seed;
  g.chk_bf = xxh3_bitflip(seed);
  return g;
}

static std::vector<uint64_t> column_bitflips(const Geo& g) {
  std::vector<uint64_t> bf(g.C);
  for (uint32_t c = 0; c < g.C; ++c) bf[c] = xxh3_bitflip(g.seed + (uint64_t)c * 5);  // column_seed, sketch.h:175
  return bf;
}

static void check_gamma(int gb) {
  if (gb != 4 && gb != 8) throw GzError(GZ_ERR_INVALID_ARG, "gamma_bytes must be 4 or 8");
}

// ================================================================================================
// GzSketch
// ================================================================================================
GzSketch::GzSketch(uint64_t vector_len, uint64_t seed, size_t samples, size_t cols, int gamma_bytes) : gb(gamma_bytes) {
  check_gamma(gb);
  require_device();
  geo = make_geo(1, vector_len, seed, samples, cols, gb);
  if (update_tiles_lds_bytes(geo, gb) > 64 * 1024) throw GzError(GZ_ERR_INVALID_ARG, "sketch too large for one workgroup's LDS");
  GZ_HIP(hipGetDevice(&dev));
  GZ_HIP(hipStreamCreate(&st));
  slab.alloc(geo.row);
  out.alloc(4);
  auto bf = column_bitflips(geo);
  col_bf.alloc(bf.size());
  GZ_HIP(hipMemcpyAsync(col_bf.p, bf.data(), bf.size() * 8, hipMemcpyHostToDevice, st));
  zero();
}
GzSketch::~GzSketch() {
  if (st) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
}
void GzSketch::zero() {  // src/sketch.cpp:86-92
  launch_fill_u32(reinterpret_cast<uint32_t*>(slab.p), 0, geo.row / 4, st);
  GZ_HIP(hipStreamSynchronize(st));
  sample_idx = 0;
}
void GzSketch::update_batch(const uint64_t* idx, size_t count) {  // src/sketch.cpp:69-83
  if (!count) return;
  constexpr size_t kTile = 1024;
  raw.ensure(count);
  const size_t nt = (count + kTile - 1) / kTile;
  std::vector<Tile> ht(nt);
  for (size_t t = 0; t < nt; ++t)
    ht[t] = Tile{0, (uint32_t)(t * kTile), (uint32_t)std::min(kTile, count - t * kTile), nt > 1 ? 1u : 0u};
  tiles.ensure(nt);
  GZ_HIP(hipMemcpyAsync(raw.p, idx, count * 8, hipMemcpyHostToDevice, st));
  GZ_HIP(hipMemcpyAsync(tiles.p, ht.data(), nt * sizeof(Tile), hipMemcpyHostToDevice, st));
  launch_update_tiles(gb, tiles.p, nullptr, (uint32_t)nt, 0, nullptr, raw.p, slab.p, geo, col_bf.p, st);
  GZ_HIP(hipGetLastError());
  GZ_HIP(hipStreamSynchronize(st));
}
int GzSketch::sample(uint64_t* idx) {  // src/sketch.cpp:94-116
  if (sample_idx >= geo.samples) throw GzError(GZ_ERR_OUT_OF_SAMPLES, "sketch cannot be sampled more times");
  const uint32_t s = (uint32_t)sample_idx++;
  launch_sketch_sample(gb, slab.p, geo, s, out.p, st);
  uint32_t h[3];
  GZ_HIP(hipMemcpyAsync(h, out.p, 12, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
  *idx = ((uint64_t)h[2] << 32) | h[1];
  return (int)h[0];
}
void GzSketch::merge(const GzSketch& o) {  // src/sketch.cpp:149-154
  if (o.geo.nb != geo.nb || o.gb != gb) throw GzError(GZ_ERR_INVALID_ARG, "sketch shapes differ");
  GZ_HIP(hipStreamSynchronize(o.st));
  launch_xor_range(gb, slab.p, o.slab.p, geo, 1, 0, geo.nb - 1, true, st);
  GZ_HIP(hipStreamSynchronize(st));
}
void GzSketch::range_merge(const GzSketch& o, size_t start, size_t ns) {  // src/sketch.cpp:156-179
  if (o.geo.nb != geo.nb || o.gb != gb) throw GzError(GZ_ERR_INVALID_ARG, "sketch shapes differ");
  if (start + ns > geo.samples) {
    sample_idx = geo.samples;  // the reference leaves the sketch in a fail state
    return;
  }
  sample_idx = std::max(sample_idx, start);
  GZ_HIP(hipStreamSynchronize(o.st));
  const uint32_t b0 = (uint32_t)(start * geo.cps * geo.D), cnt = (uint32_t)(ns * geo.cps * geo.D);
  launch_xor_range(gb, slab.p, o.slab.p, geo, 1, b0, cnt, true, st);
  GZ_HIP(hipStreamSynchronize(st));
}
void GzSketch::merge_raw(const void* rawb) {  // src/sketch.cpp:181-186
  packed.ensure(bytes());
  GZ_HIP(hipMemcpyAsync(packed.p, rawb, bytes(), hipMemcpyHostToDevice, st));
  launch_merge_packed(gb, slab.p, geo, 0, 1, packed.p, st);
  GZ_HIP(hipStreamSynchronize(st));
}
void GzSketch::serialize(void* o, size_t nbytes) const {  // src/sketch.cpp:188-190
  if (nbytes < bytes()) throw GzError(GZ_ERR_INVALID_ARG, "output buffer smaller than bucket_array_bytes()");
  auto* self = const_cast<GzSketch*>(this);
  self->packed.ensure(bytes());
  launch_pack(gb, slab.p, geo, 0, 1, self->packed.p, st);
  GZ_HIP(hipMemcpyAsync(o, packed.p, bytes(), hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
}
void GzSketch::deserialize(const void* in, size_t nbytes) {  // src/sketch.cpp:23-35
  if (nbytes < bytes()) throw GzError(GZ_ERR_INVALID_ARG, "input buffer smaller than bucket_array_bytes()");
  packed.ensure(bytes());
  GZ_HIP(hipMemcpyAsync(packed.p, in, bytes(), hipMemcpyHostToDevice, st));
  launch_unpack(gb, slab.p, geo, 0, 1, packed.p, st);
  GZ_HIP(hipStreamSynchronize(st));
}

// ================================================================================================
// GzAlg
// ================================================================================================
static size_t default_cap_pairs(uint32_t n, size_t desired) {
  const size_t want = (size_t)n * desired / 2;
  return std::min<size_t>(size_t(1) << 25, std::max<size_t>(size_t(1) << 20, want));
}

uint64_t GzAlg::footprint(uint32_t n, const gz_alg_config& c) {
  const uint64_t len = calc_vector_length(n);
  const size_t D = calc_bkt_per_col(len), C = calc_cc_samples(n, c.sketches_factor);
  const uint64_t nb = C * D + 1;
  const int gbytes = c.gamma_bytes ? c.gamma_bytes : 4;
  const uint64_t slab = (uint64_t)n * row_layout((uint32_t)C, (uint32_t)D, 8 + gbytes).row;
  const size_t desired = (size_t)(nb * (8 + gbytes) / 4 * c.batch_factor);
  const uint64_t cap = c.flush_threshold ? c.flush_threshold : default_cap_pairs(n, desired);
  const uint32_t t = c.tile ? c.tile : 1024;
  const uint64_t part = n <= (1u << 24) ? cap * 16 /*records*/ + 8 * 1048577ull + (1 << 16) : 0;
  const uint64_t ingest = cap * 8 /*pending*/ + cap * 8 /*dst_arr*/ + cap /*merged*/ + part +
                          (2 * cap / t + n + 1) * sizeof(Tile) + (uint64_t)n * 20;
  // per vertex: cnt, off, cursor, 2 member lists, 2 member-root lists, parent, labels (4 B each),
  // sample record, forest edge, new root and shift (8 each), 2 segment lists (16 each), spanning-forest
  // hash set (16), deterministic-bucket copy, degree class (1); the reduce's cross-workgroup
  // accumulators (1 KiB + 4 B per 1024 vertices, and one more slot)
  const uint64_t rounds =
      (uint64_t)n * (9 * 4 + 8 + 8 + 8 + 2 * 16 + 16 + 8 + gbytes + 1) + ((uint64_t)n / 512 + 3) * 1028 + 8192;
  return slab + ingest + rounds;
}

GzAlg::GzAlg(uint32_t n_, uint64_t seed_, const gz_alg_config& c) : n(n_), seed(seed_), cfg(c) {
  if (n < 2) throw GzError(GZ_ERR_INVALID_ARG, "num_vertices must be >= 2");
  if (n > (1u << 31)) throw GzError(GZ_ERR_INVALID_ARG, "num_vertices above 2^31 is not supported");
  if (cfg.sketches_factor <= 0) cfg.sketches_factor = 1;  // cc_alg_configuration.cpp:15-23
  if (cfg.batch_factor <= 0) cfg.batch_factor = 1;        // cc_alg_configuration.cpp:25-33
  if (cfg.gamma_bytes == 0) cfg.gamma_bytes = 4;
  gb = cfg.gamma_bytes;
  check_gamma(gb);
  require_device();
  geo = make_geo(n, calc_vector_length(n), seed, calc_cc_samples(n, cfg.sketches_factor), 1, gb);
  if (update_tiles_lds_bytes(geo, gb) > 64 * 1024) throw GzError(GZ_ERR_INVALID_ARG, "sketch too large for LDS");
  tile = cfg.tile ? cfg.tile : 1024;
  const size_t lds = update_tiles_lds_bytes(geo, gb);
  // one-wave workgroups: up to 32 per CU (8 waves on each of 4 SIMDs), fewer if LDS runs out
  const size_t per_cu = std::max<size_t>(1, std::min<size_t>(32, (160 * 1024) / lds));
  grid_cap = (uint32_t)(device_cu_count() * per_cu);
  GZ_HIP(hipStreamCreate(&st));

  // sketch slab
  slab.alloc((size_t)n * geo.row);
  launch_fill_u32(reinterpret_cast<uint32_t*>(slab.p), 0, slab.n / 4, st);
  auto bf = column_bitflips(geo);
  col_bf.alloc(bf.size());
  GZ_HIP(hipMemcpyAsync(col_bf.p, bf.data(), bf.size() * 8, hipMemcpyHostToDevice, st));

  // DSU + forest
  parent.alloc(n);
  uint64_t cap = 1;
  while (cap < 2ull * n) cap <<= 1;
  sf_hash.alloc(cap);
  sf_edge.alloc(n);
  flags.alloc(kNumFlags);
  mod.alloc(geo.samples + 2);
  labels.alloc(n + 1);
  dsu.parent = parent.p;
  dsu.sf_hash = sf_hash.p;
  dsu.sf_mask = cap - 1;
  dsu.sf_edge = sf_edge.p;
  dsu.flags = flags.p;
  dsu.mod = mod.p;
  dsu.nmod = geo.samples + 2;
  dsu.prio_seed = (uint32_t)(seed * 0x9E3779B97F4A7C15ULL >> 32);
  dsu.n = n;
  eager_inv.ensure(1);
  eager_inv.p[0] = 0;
  dsu.inv_mirror = eager_inv.d;
  qforce.ensure(1);
  qforce.p[0] = 1;
  dsu.query_force = qforce.d;
  launch_fill_u32(flags.p, 0, kNumFlags, st);  // the reset below keeps the bad-input flag
  launch_dsu_reset(dsu, st);

  // Borůvka round buffers
  r_cnt.alloc(n); r_off.alloc(n); r_cursor.alloc(n); r_samp.alloc(n);
  r_members[0].alloc(n); r_members[1].alloc(n); r_mroot[0].alloc(n); r_mroot[1].alloc(n);
  r_segs[0].alloc(n); r_segs[1].alloc(n); r_nd.alloc(n);
  r_totals.alloc(16);
  r_stats.alloc(128);
  r_det.alloc((size_t)n * (8 + gb));
  r_hint.alloc(n);
  launch_fill_u32(r_stats.p, 0, 128, st);
  // cross-workgroup accumulators of the reduce: one slot per workgroup and one for the giant
  const size_t xslots = (size_t)reduce_blocks(n) + 1;
  r_xacc.alloc(xslots * 128);
  r_xcnt.alloc(xslots);
  launch_fill_u64(r_xacc.p, 0, xslots * 128, st);
  launch_fill_u32(r_xcnt.p, 0, xslots, st);
  rb = RoundBufs{r_cnt.p, r_off.p, r_cursor.p, {r_members[0].p, r_members[1].p},
                 {r_mroot[0].p, r_mroot[1].p}, {r_segs[0].p, r_segs[1].p}, r_nd.p, r_totals.p,
                 r_xacc.p, r_xcnt.p, (uint32_t)xslots, r_samp.p, r_stats.p, r_det.p, r_hint.p};
  launch_fill_u32(r_cnt.p, 0, n, st);
  // Round kernels read the state words and the previous member list before their round flag, so
  // both hold in-range values from the start (empty walks, vertex 0); every query's first kernel
  // sets the state words again.
  launch_round_state_init(rb, n, st);

  // device ingest
  cap_pairs = cfg.flush_threshold ? (size_t)cfg.flush_threshold : default_cap_pairs(n, desired_updates_per_batch());
  if (cap_pairs > (size_t(1) << 30)) cap_pairs = size_t(1) << 30;
  pending.alloc(2 * cap_pairs);
  dst_arr.alloc(2 * cap_pairs);
  merged.alloc(cap_pairs);
  g_cnt.alloc(n); g_off.alloc(n + 1); g_toff.alloc(n + 1); g_cursor.alloc(n); g_totals.alloc(4);
  g_partials.alloc(group_partials_len(n));
  g_tiles.alloc(2 * cap_pairs / tile + n + 1);
  gbufs = GroupBufs{g_cnt.p, g_off.p, g_toff.p, g_cursor.p, g_partials.p, g_totals.p, dst_arr.p, g_tiles.p,
                    flags.p + kFlagBadInput, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (n <= (1u << 24)) {  // bucket-partition grouping (no device-scope atomics)
    g_rec.alloc(2 * cap_pairs);
    g_hist.alloc(part_hist_len());
    g_hoff.alloc(part_hist_len());
    g_hpartials.alloc(group_partials_len((uint32_t)part_hist_len()));
    g_htotals.alloc(4);
    gbufs.rec = g_rec.p;
    gbufs.hist = g_hist.p;
    gbufs.hoff = g_hoff.p;
    gbufs.hpartials = g_hpartials.p;
    gbufs.htotals = g_htotals.p;
  }

  // host staging
  stage_cap = std::max<size_t>(desired_updates_per_batch(), cfg.batched ? (size_t(1) << 22) : 0);
  stages.resize(cfg.batched ? 2 : 4);
  vstamp.assign(n, 0);
  vtiles.assign(n, 0);
  eager_cap = size_t(1) << 20;
  hflags.alloc(std::max<size_t>(kNumFlags, geo.samples + 2));
  hsmall.alloc(16);
  hround.alloc(128);
  rep.ensure(dsu.nmod + 128);
  GZ_HIP(hipEventCreateWithFlags(&hround_ev, hipEventDisableTiming));
  GZ_HIP(hipEventCreateWithFlags(&eager_ev, hipEventDisableTiming));
  GZ_HIP(hipStreamSynchronize(st));
}

GzAlg::~GzAlg() {
  if (st) (void)hipStreamSynchronize(st);
  if (cst) {
    (void)hipStreamSynchronize(cst);
    (void)hipStreamDestroy(cst);
  }
  if (copy_ev) (void)hipEventDestroy(copy_ev);
  if (pend_free_ev) (void)hipEventDestroy(pend_free_ev);
  for (auto& h : hstage)
    if (h.ev) (void)hipEventDestroy(h.ev);
  for (auto& t : timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  for (auto& s : stages)
    if (s.ev) (void)hipEventDestroy(s.ev);
  if (eager_ev) (void)hipEventDestroy(eager_ev);
  if (hround_ev) (void)hipEventDestroy(hround_ev);
  for (auto& kv : query_graphs) (void)hipGraphExecDestroy(kv.second);
  if (st) (void)hipStreamDestroy(st);
}

// ---- ingest --------------------------------------------------------------------------------------

// pre_insert, src/cc_sketch_alg.cpp:79-104: buffered and applied on the device in batches.
void GzAlg::pre_insert(uint32_t src, uint32_t dst) {
  if (src >= n || dst >= n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
  if (!cfg.eager_dsu) {
    if (dsu_valid) dsu_valid = false;
    return;
  }
  if (!dsu_valid) return;
  std::lock_guard<std::mutex> lk(eager_mtx);
  if (!eager_host.p) {
    eager_host.alloc(2 * eager_cap);
    eager_dev.alloc(2 * eager_cap);
    eager_merged.alloc(eager_cap);
  }
  if (eager_used == eager_cap) submit_eager_locked();
  eager_host.p[2 * eager_used] = src;
  eager_host.p[2 * eager_used + 1] = dst;
  ++eager_used;
}

void GzAlg::submit_eager_locked() {
  if (!eager_used) return;
  GZ_HIP(hipMemcpyAsync(eager_dev.p, eager_host.p, eager_used * 8, hipMemcpyHostToDevice, st));
  eager_batch(eager_dev.p, eager_dev.p, eager_used, eager_merged.p);
  GZ_HIP(hipGetLastError());
  GZ_HIP(hipEventRecord(eager_ev, st));
  GZ_HIP(hipEventSynchronize(eager_ev));  // the host buffer is reused right away
  eager_used = 0;
}

// apply_update_batch, src/cc_sketch_alg.cpp:106-118
void GzAlg::apply_update_batch(uint32_t src, const uint32_t* dsts, size_t k) {
  if (update_locked) throw GzError(GZ_ERR_UPDATE_LOCKED, "Cannot update the algorithm: Connected components currently running");
  if (src >= n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
  if (!k) return;
  if (k > 0xffffffffull) throw GzError(GZ_ERR_INVALID_ARG, "batch too large");
  std::lock_guard<std::mutex> lk(stage_mtx);
  zero_sum = 0;  // one direction of an edge alone would break the zero sum
  if (!cfg.batched) {
    plain_launch(src, dsts, k);
    return;
  }
  Stage* s = &stages[cur_stage];
  if (s->used + k > s->dcap || s->ent_v.size() >= kMaxEntries) {
    submit_stage_locked();
    s = &stages[cur_stage];
  }
  if (s->inflight) {
    GZ_HIP(hipEventSynchronize(s->ev));
    s->inflight = false;
  }
  if (s->dcap < k) {  // first use, or a batch larger than the stage: (re)allocate an empty stage
    const size_t dcap = std::max(stage_cap, k);
    s->max_tiles = dcap / tile + kMaxEntries + 1;
    s->host.alloc(s->max_tiles * sizeof(Tile) + dcap * 4);
    s->dev.alloc(s->host.n);
    s->dcap = dcap;
  }
  uint32_t* hd = reinterpret_cast<uint32_t*>(s->host.p + s->max_tiles * sizeof(Tile));
  std::memcpy(hd + s->used, dsts, k * 4);
  s->ent_v.push_back(src);
  s->ent_off.push_back((uint32_t)s->used);
  s->ent_cnt.push_back((uint32_t)k);
  s->used += k;
  total_updates += k;
}

// Submit the current batched stage: build tiles (split flag when a vertex has several tiles in
// this launch), copy tiles and destinations to the device and launch one update kernel over all.
// Host layout of a stage: [max_tiles Tile records | dcap destination ids].
void GzAlg::submit_stage_locked() {
  Stage& s = stages[cur_stage];
  if (!s.used) return;
  ++epoch;
  if (epoch == 0) {
    std::fill(vstamp.begin(), vstamp.end(), 0u);
    epoch = 1;
  }
  for (size_t e = 0; e < s.ent_v.size(); ++e) {
    const uint32_t v = s.ent_v[e];
    if (vstamp[v] != epoch) {
      vstamp[v] = epoch;
      vtiles[v] = 0;
    }
    vtiles[v] += (s.ent_cnt[e] + tile - 1) / tile;
  }
  Tile* ht = reinterpret_cast<Tile*>(s.host.p);
  size_t nt = 0;
  for (size_t e = 0; e < s.ent_v.size(); ++e) {
    const uint32_t v = s.ent_v[e], c = s.ent_cnt[e], o = s.ent_off[e];
    const uint32_t split = vtiles[v] > 1 ? 1u : 0u;
    for (uint32_t b = 0; b < c; b += tile) ht[nt++] = Tile{v, o + b, std::min(tile, c - b), split};
  }
  const size_t doff = s.max_tiles * sizeof(Tile);
  GZ_HIP(hipMemcpyAsync(s.dev.p, s.host.p, nt * sizeof(Tile), hipMemcpyHostToDevice, st));
  GZ_HIP(hipMemcpyAsync(s.dev.p + doff, s.host.p + doff, s.used * 4, hipMemcpyHostToDevice, st));
  if (prof) prof_begin(0, batch_alg_bytes(s.ent_cnt), false, s.used);
  launch_update_tiles(gb, reinterpret_cast<Tile*>(s.dev.p), nullptr, (uint32_t)nt, 0,
                      reinterpret_cast<uint32_t*>(s.dev.p + doff), nullptr, slab.p, geo, col_bf.p, st);
  if (prof) prof_end();
  GZ_HIP(hipGetLastError());
  if (!s.ev) GZ_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  GZ_HIP(hipEventRecord(s.ev, st));
  s.inflight = true;
  s.used = 0;
  s.ent_v.clear();
  s.ent_off.clear();
  s.ent_cnt.clear();
  cur_stage = (cur_stage + 1) % stages.size();
}

double GzAlg::batch_alg_bytes(const std::vector<uint32_t>& counts) const {
  const double S = (double)sketch_bytes(), per = (double)(geo.C + 1) * (8 + gb);
  double b = 0;
  for (uint32_t k : counts) b += 4.0 * k + 2.0 * std::min(S, k * per);
  return b;
}

// Plain form: one launch per apply_update_batch call (one gutter flush).
void GzAlg::plain_launch(uint32_t src, const uint32_t* dsts, size_t k) {
  Stage& s = stages[cur_stage];
  cur_stage = (cur_stage + 1) % stages.size();
  if (s.inflight) {
    GZ_HIP(hipEventSynchronize(s.ev));
    s.inflight = false;
  }
  const size_t nt = (k + tile - 1) / tile;
  const size_t tbytes = ((nt * sizeof(Tile)) + 255) & ~size_t(255);
  const size_t need = tbytes + k * 4;
  if (s.host.n < need) {
    s.host.alloc(need);
    s.dev.alloc(need);
  }
  Tile* ht = reinterpret_cast<Tile*>(s.host.p);
  for (size_t t = 0; t < nt; ++t)
    ht[t] = Tile{src, (uint32_t)(t * tile), (uint32_t)std::min<size_t>(tile, k - t * tile), nt > 1 ? 1u : 0u};
  std::memcpy(s.host.p + tbytes, dsts, k * 4);
  GZ_HIP(hipMemcpyAsync(s.dev.p, s.host.p, need, hipMemcpyHostToDevice, st));
  if (prof) prof_begin(0, batch_alg_bytes(std::vector<uint32_t>{(uint32_t)k}), false, k);
  launch_update_tiles(gb, reinterpret_cast<Tile*>(s.dev.p), nullptr, (uint32_t)nt, 0,
                      reinterpret_cast<uint32_t*>(s.dev.p + tbytes), nullptr, slab.p, geo, col_bf.p, st);
  if (prof) prof_end();
  GZ_HIP(hipGetLastError());
  if (!s.ev) GZ_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  GZ_HIP(hipEventRecord(s.ev, st));
  s.inflight = true;
  total_updates += k;
}

// apply_raw_buckets_update, src/cc_sketch_alg.cpp:120-123
void GzAlg::apply_raw(uint32_t src, const void* rawb) {
  if (src >= n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
  std::lock_guard<std::mutex> lk(stage_mtx);
  DevBuf<uint8_t> tmp;
  tmp.alloc(sketch_bytes());
  GZ_HIP(hipMemcpyAsync(tmp.p, rawb, sketch_bytes(), hipMemcpyHostToDevice, st));
  launch_merge_packed(gb, slab.p, geo, src, 1, tmp.p, st);
  GZ_HIP(hipStreamSynchronize(st));
  zero_sum = 0;  // a raw delta on one vertex may break the zero sum
}

// ---- kernel timing ----
hipEvent_t GzAlg::take_event() {
  if (!ev_pool.empty()) {
    hipEvent_t e = ev_pool.back();
    ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  GZ_HIP(hipEventCreate(&e));
  return e;
}
hipEvent_t GzAlg::prof_begin(int kind, double host_bytes, bool dev_bytes, uint64_t items) {
  Timed t{take_event(), take_event(), kind, host_bytes, dev_bytes, items};
  GZ_HIP(hipEventRecord(t.a, st));
  timed.push_back(t);
  return t.a;
}
void GzAlg::prof_end() { GZ_HIP(hipEventRecord(timed.back().b, st)); }
// Called after the stream has been synchronised: folds finished event pairs into the profile.
void GzAlg::prof_collect() {
  if (timed.empty()) return;
  for (auto& t : timed) {
    float ms = 0;
    GZ_HIP(hipEventSynchronize(t.b));
    GZ_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    gz_kernel_stat& k = t.kind == 0 ? profile.update : profile.boruvka;
    k.launches += 1;
    k.total_ms += ms;
    k.alg_bytes += t.host_bytes;
    k.items += t.items;
    ev_pool.push_back(t.a);
    ev_pool.push_back(t.b);
  }
  timed.clear();
  std::vector<uint64_t> slots(kProfSlots);
  GZ_HIP(hipMemcpy(slots.data(), prof_bytes.p, kProfSlots * 8, hipMemcpyDeviceToHost));
  uint64_t db = 0;
  for (uint64_t x : slots) db += x;
  profile.update.alg_bytes += (double)db;
  launch_fill_u64(prof_bytes.p, 0, kProfSlots, st);
  GZ_HIP(hipStreamSynchronize(st));
}

void GzAlg::flush_pending() {
  const uint32_t* src = ref_n ? ref_pairs : pending.p;
  const size_t cnt = ref_n ? ref_n : pending_n;
  if (!cnt) return;
  const FlushBytes fb{prof_bytes.p, sketch_bytes(), (uint64_t)(geo.C + 1) * (8 + gb)};
  launch_group(src, cnt, n, tile, gbufs, st, prof ? &fb : nullptr);
  if (!ref_n && cst) {  // host copies into `pending` on the copy stream wait until the grouping read it
    GZ_HIP(hipEventRecord(pend_free_ev, st));
    pend_wait = true;
  }
  if (prof) prof_begin(0, 0, true, 2 * cnt);
  launch_update_tiles(gb, g_tiles.p, g_totals.p + 1, 0, grid_cap, dst_arr.p, nullptr, slab.p,
                      geo, col_bf.p, st);
  if (prof) prof_end();
  GZ_HIP(hipGetLastError());
  total_updates += 2 * cnt;
  pending_n = 0;
  ref_pairs = nullptr;
  ref_n = 0;
}

void GzAlg::ingest_device(const uint32_t* pairs_dev, size_t count) {
  if (update_locked) throw GzError(GZ_ERR_UPDATE_LOCKED, "Cannot update the algorithm: Connected components currently running");
  if (!count) return;
  std::lock_guard<std::mutex> lk(stage_mtx);
  if (!cfg.eager_dsu) dsu_valid = false;
  if (ref_n) flush_pending();  // referenced pairs and copied pairs are never mixed in one flush
  while (count) {
    const size_t take = std::min(count, cap_pairs - pending_n);
    uint32_t* dst = pending.p + 2 * pending_n;
    if (cfg.eager_dsu && dsu_valid && eager_live()) eager_batch(pairs_dev, dst, take, merged.p);  // copies as well
    else GZ_HIP(hipMemcpyAsync(dst, pairs_dev, take * 8, hipMemcpyDeviceToDevice, st));
    pending_n += take;
    pairs_dev += 2 * take;
    count -= take;
    if (pending_n == cap_pairs) flush_pending();
  }
}

void GzAlg::ingest_device_stable(const uint32_t* pairs_dev, size_t count) {
  if (update_locked) throw GzError(GZ_ERR_UPDATE_LOCKED, "Cannot update the algorithm: Connected components currently running");
  if (!count) return;
  std::lock_guard<std::mutex> lk(stage_mtx);
  if (!cfg.eager_dsu) dsu_valid = false;
  if (pending_n || (ref_n && pairs_dev != ref_pairs + 2 * ref_n)) flush_pending();
  while (count) {
    const size_t take = std::min(count, cap_pairs - ref_n);
    // the eager batch reads the pairs in place (no copy); once the eager DSU is known to be invalid
    // (the mapped mirror of its flag) no eager launch is needed until the next query
    if (cfg.eager_dsu && dsu_valid && eager_live()) eager_batch(pairs_dev, const_cast<uint32_t*>(pairs_dev), take, merged.p);
    if (!ref_n) ref_pairs = pairs_dev;
    ref_n += take;
    pairs_dev += 2 * take;
    count -= take;
    if (ref_n == cap_pairs) flush_pending();
  }
}

void GzAlg::release_stable() {
  std::lock_guard<std::mutex> lk(stage_mtx);
  if (cst) (void)hipStreamSynchronize(cst);  // host pairs read in place have landed
  if (!ref_n) return;
  try {
    flush_pending();
    GZ_HIP(hipStreamSynchronize(st));
  } catch (...) {
    ref_pairs = nullptr;
    ref_n = 0;
    throw;
  }
}

void GzAlg::ensure_host_ingest() {
  if (cst) return;
  GZ_HIP(hipStreamCreate(&cst));
  GZ_HIP(hipEventCreateWithFlags(&copy_ev, hipEventDisableTiming));
  GZ_HIP(hipEventCreateWithFlags(&pend_free_ev, hipEventDisableTiming));
  for (auto& h : hstage) {
    h.buf.alloc(2 * kHostStagePairs);
    GZ_HIP(hipEventCreateWithFlags(&h.ev, hipEventDisableTiming));
  }
  // work already on the compute stream may still read `pending`
  GZ_HIP(hipEventRecord(pend_free_ev, st));
  pend_wait = true;
}

// Whether p lies in page-locked host memory the device can read directly.  Pageable memory either
// reports an unregistered type or an error, which is cleared here.
static bool host_pinned(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

void GzAlg::ingest_host(const uint32_t* pairs, size_t count, bool stable) {
  if (update_locked) throw GzError(GZ_ERR_UPDATE_LOCKED, "Cannot update the algorithm: Connected components currently running");
  if (!count) return;
  std::lock_guard<std::mutex> lk(stage_mtx);
  if (!cfg.eager_dsu) dsu_valid = false;
  if (ref_n) flush_pending();
  ensure_host_ingest();
  const bool direct = host_pinned(pairs);
  while (count) {
    size_t take = std::min(count, cap_pairs - pending_n);
    const void* src = pairs;
    HostStage* hs = nullptr;
    if (!direct) {
      take = std::min(take, kHostStagePairs);
      hs = &hstage[hstage_cur];
      hstage_cur ^= 1;
      if (hs->busy) GZ_HIP(hipEventSynchronize(hs->ev));  // its previous copy has landed
      std::memcpy(hs->buf.p, pairs, take * 8);
      src = hs->buf.p;
    }
    uint32_t* dst = pending.p + 2 * pending_n;
    if (pend_wait) {
      GZ_HIP(hipStreamWaitEvent(cst, pend_free_ev, 0));
      pend_wait = false;
    }
    GZ_HIP(hipMemcpyAsync(dst, src, take * 8, hipMemcpyHostToDevice, cst));
    GZ_HIP(hipEventRecord(copy_ev, cst));
    if (hs) {
      GZ_HIP(hipEventRecord(hs->ev, cst));
      hs->busy = true;
    }
    GZ_HIP(hipStreamWaitEvent(st, copy_ev, 0));  // the kernels that read `dst` follow the copy
    if (cfg.eager_dsu && dsu_valid && eager_live()) eager_batch(dst, dst, take, merged.p);
    pending_n += take;
    pairs += 2 * take;
    count -= take;
    if (pending_n == cap_pairs) flush_pending();
  }
  // pairs read in place are only borrowed for the call
  if (direct && !stable) GZ_HIP(hipEventSynchronize(copy_ev));
}

// The driver's force_flush + flush_workers (include/graph_sketch_driver.h:193-203)
void GzAlg::flush_all() {
  {
    std::lock_guard<std::mutex> lk(stage_mtx);
    if (cfg.batched) submit_stage_locked();
    flush_pending();
  }
  GZ_HIP(hipStreamSynchronize(st));
  if (prof) prof_collect();
  if (read_flag(kFlagBadInput)) {
    launch_fill_u32(flags.p + kFlagBadInput, 0, 1, st);
    GZ_HIP(hipStreamSynchronize(st));
    throw GzError(GZ_ERR_INVALID_ARG, "ingested device pairs named a vertex >= num_vertices; those pairs were skipped");
  }
}

uint32_t GzAlg::read_flag(int slot) {
  GZ_HIP(hipMemcpyAsync(hflags.p, flags.p, kNumFlags * 4, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
  return hflags.p[slot];
}

void GzAlg::sync_eager_state() {
  if (!cfg.eager_dsu) return;
  {
    std::lock_guard<std::mutex> lk(eager_mtx);
    if (dsu_valid) submit_eager_locked();
    else eager_used = 0;
  }
  if (dsu_valid && read_flag(kFlagEagerInvalid)) dsu_valid = false;
}

// has_cached_query, include/cc_sketch_alg.h:184-190
bool GzAlg::has_cached_query(int code) {
  if (code != GZ_CONNECTIVITY) return false;
  sync_eager_state();
  return dsu_valid;
}

void GzAlg::prep_query(int code) {
  if (has_cached_query(code)) return;
  flush_all();
}

void GzAlg::ensure_answer() {
  sync_eager_state();
  if (dsu_valid) return;
  flush_all();
  boruvka();
}

// Edge updates keep the XOR of all vertex sketches at zero; raw bucket deltas and loaded images
// may not, so after those the sum is checked once (one pass over the sketches).
bool GzAlg::sketches_sum_to_zero() {
  if (zero_sum == 0) {
    DevBuf<uint64_t> sa, sg;
    sa.alloc(geo.nb);
    sg.alloc(geo.nb);
    launch_sum_all(gb, slab.p, geo, sa.p, sg.p, st);
    std::vector<uint64_t> ha(geo.nb), hg(geo.nb);
    GZ_HIP(hipMemcpyAsync(ha.data(), sa.p, geo.nb * 8, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipMemcpyAsync(hg.data(), sg.p, geo.nb * (size_t)gb, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    bool zero = true;
    for (uint32_t b = 0; b < geo.nb && zero; ++b) zero = ha[b] == 0;
    const uint8_t* gbytes = reinterpret_cast<const uint8_t*>(hg.data());
    for (size_t i = 0; i < geo.nb * (size_t)gb && zero; ++i) zero = gbytes[i] == 0;
    zero_sum = zero ? 1 : -1;
  }
  return zero_sum == 1;
}

// DSU reset, round 0, rounds 1..last and (with a giant allowed) the single-workgroup finish from
// round last + 1, as one stream sequence (captured into a graph or launched directly).  Fills are
// kernels, so the sequence holds no memset nodes.
void GzAlg::enqueue_query_prefix(bool giant_ok, uint32_t last) {
  launch_dsu_reset(dsu, st, &rb);  // with the root counts and the round state words
  launch_sample_round0(gb, slab.p, geo, dsu, rb, st);
  for (uint32_t r = 1; r <= last; ++r) launch_round(gb, r, slab.p, geo, dsu, rb, giant_ok, st, use_deep_skip);
  if (giant_ok && use_finish) launch_finish(gb, last + 1, geo.samples - 1, slab.p, geo, dsu, rb, st, rep.d);
}

hipGraphExec_t GzAlg::get_query_graph(bool giant_ok, uint32_t last) {
  const uint32_t key = ((last * 2 + (giant_ok ? 1 : 0)) * 2 + (use_finish ? 1 : 0)) * 2 + (use_deep_skip ? 1 : 0);
  hipGraphExec_t& slot = query_graphs[key];
  if (slot) return slot;
  GZ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  enqueue_query_prefix(giant_ok, last);
  hipGraph_t graph = nullptr;
  const hipError_t cap = hipStreamEndCapture(st, &graph);
  if (cap != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    query_graphs.erase(key);
    throw GzError(GZ_ERR_DEVICE, std::string("query graph capture: ") + hipGetErrorString(cap));
  }
  const hipError_t inst = hipGraphInstantiate(&slot, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (inst != hipSuccess) {
    query_graphs.erase(key);
    throw GzError(GZ_ERR_DEVICE, std::string("query graph instantiate: ") + hipGetErrorString(inst));
  }
  return slot;
}

// boruvka_emulation, src/cc_sketch_alg.cpp:464-513.  Every kernel of round r is a no-op once
// round r-1 changed nothing, so rounds are queued ahead and the host reads the per-round flags
// once per group: first the graph prefix (round 0, rounds 1..K and a finish attempt), then groups
// of 2 rounds, each followed by a finish attempt.  A finish that takes over runs every remaining
// round, so at most one does per query.
bool GzAlg::boruvka(const std::function<void()>* speculate, bool gated, bool* skipped) {
  bool answered = false;
  if (skipped) *skipped = false;
  gated = gated && speculate;  // the skip is read from the speculated answers' flags
  qforce.p[0] = gated ? 0u : 1u;  // read by the query's first kernel, after this store
  update_locked = true;
  try {
    const bool giant_ok = sketches_sum_to_zero();
    const uint32_t last_round = geo.samples - 1;  // round index `samples` would be out of samples
    if (hround_pending) {
      // The launched rounds of this query: up to the first round of the previous query after which
      // a giant existed and at most finish_max() vertices remained outside it (diagnostics: vertices
      // regrouped and the giant's root per round).
      if (hround_on_event) GZ_HIP(hipEventSynchronize(hround_ev));
      hround_pending = false;
      next_prefix = (uint32_t)last_rounds + 1;
      for (uint32_t r = 1; r < (uint32_t)last_rounds && r < 64; ++r)
        if (hround.p[2 * r + 1] != 0xFFFFFFFFu && hround.p[2 * r] <= finish_max()) {
          next_prefix = r;
          break;
        }
    }
    const uint32_t want = next_prefix ? next_prefix : kFirstGraphRounds;
    const uint32_t first = std::max(1u, std::min(want, last_round));
    const bool sf_was_pending = sf_pending;
    sf_pending = false;  // the DSU reset clears the set
    if (prof) prof_begin(1, 0, false, 0);
    if (use_graphs) GZ_HIP(hipGraphLaunch(get_query_graph(giant_ok, first), st));
    else enqueue_query_prefix(giant_ok, first);
    GZ_HIP(hipGetLastError());
    uint32_t next = first + 1;
    uint32_t fin_r0 = first + 1;  // first round of the last finish attempt queued
    int64_t rounds = -1;
    bool first_group = true;
    // A group that ends with the finish kernel gets its round flags (and the diagnostics) from the
    // kernel's report in mapped memory; otherwise they are copied back.
    bool reported = giant_ok && use_finish && first + 1 >= 2 && first + 1 <= last_round;
    const uint32_t* hf = hflags.p;
    while (true) {
      if (!reported) GZ_HIP(hipMemcpyAsync(hflags.p, mod.p, (size_t)dsu.nmod * 4, hipMemcpyDeviceToHost, st));
      if (first_group && speculate) (*speculate)();
      GZ_HIP(hipStreamSynchronize(st));
      hf = reported ? rep.p : hflags.p;
      // rounds launched so far, or as far as a finish that took over went
      const uint32_t upto = std::max(next - 1, hf[dsu.nmod - 1]);
      for (uint32_t r = 0; r <= upto; ++r)
        if (!hf[r]) {
          rounds = r;
          break;
        }
      if (first_group && gated && aq_out.p[kFlagSkip]) {
        // the eager DSU was still valid: the query's kernels kept it, and the answers came from it
        if (prof) {
          ev_pool.push_back(timed.back().a);
          ev_pool.push_back(timed.back().b);
          timed.pop_back();
        }
        if (skipped) *skipped = true;
        sf_pending = sf_was_pending;  // no reset ran: the forest set is as it was
        update_locked = false;
        return true;
      }
      if (rounds >= 0) {
        answered = first_group && speculate;
        break;
      }
      first_group = false;
      if (upto >= last_round)
        throw GzError(GZ_ERR_OUT_OF_SAMPLES, "Borůvka needs more rounds than the sketches have samples");
      const uint32_t hi = std::min(next + 1, last_round);
      for (uint32_t r = next; r <= hi; ++r) launch_round(gb, r, slab.p, geo, dsu, rb, giant_ok, st, use_deep_skip);
      reported = giant_ok && use_finish && launch_finish(gb, hi + 1, last_round, slab.p, geo, dsu, rb, st, rep.d);
      GZ_HIP(hipGetLastError());
      fin_r0 = hi + 1;
      next = hi + 1;
    }
    if (prof) prof_end();
    last_rounds = (size_t)rounds;
    last_finish_start = hf[dsu.nmod - 1] ? fin_r0 : 0;
    if (reported) {
      std::memcpy(hround.p, rep.p + dsu.nmod, 128 * 4);
      hround_on_event = false;
    } else {
      GZ_HIP(hipMemcpyAsync(hround.p, r_stats.p, 128 * 4, hipMemcpyDeviceToHost, st));
      GZ_HIP(hipEventRecord(hround_ev, st));
      hround_on_event = true;
    }
    if (prof) {
      // algorithmic bytes of the query: round 0 reads depths 0..14 of column 0 and the deterministic
      // bucket of every vertex and writes a sample record; round r >= 1 reads column r and the
      // deterministic bucket ((D+1) records) and a root and a member word of every vertex it
      // regroups (those outside a giant supernode, as recorded per round)
      if (hround_on_event) GZ_HIP(hipEventSynchronize(hround_ev));
      const double rec = 8.0 + gb;
      double bytes = (double)n * (std::min<uint32_t>(geo.D, 15) + 1) * rec + 8.0 * n;
      for (int64_t r = 1; r <= rounds && r < 64; ++r) bytes += (double)hround.p[2 * r] * ((geo.D + 1) * rec + 8.0);
      timed.back().host_bytes = bytes;
      timed.back().items = (uint64_t)rounds + 1;
    }
    hround_pending = true;
    sf_pending = cfg.eager_dsu;
    // every eager batch enqueued before the query has finished (the query ended with a stream
    // synchronisation), so none can set the mirror after this
    eager_inv.p[0] = 0;
    dsu_valid = true;
    if (prof) prof_collect();
  } catch (...) {
    update_locked = false;
    dsu_valid = false;
    throw;
  }
  update_locked = false;
  return answered;
}

void GzAlg::eager_batch(const uint32_t* src, uint32_t* pairs, size_t count, uint8_t* merged) {
  if (!count) return;
  launch_eager(src, pairs, count, merged, dsu, st, sf_pending);  // builds the forest set when pending
  sf_pending = false;
}

// connected_components, src/cc_sketch_alg.cpp:515-557 + ConnectedComponents, return_types.cpp:5-19
uint32_t GzAlg::connected_components(uint32_t* parent_out) {
  std::lock_guard<std::mutex> lk(query_mtx);
  ensure_answer();
  launch_labels(dsu, labels.p, labels.p + n, st);
  uint32_t ncc = 0;
  if (parent_out) GZ_HIP(hipMemcpyAsync(parent_out, labels.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipMemcpyAsync(hsmall.p, labels.p + n, 4, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
  ncc = hsmall.p[0];
  return ncc;
}

// point_query, src/cc_sketch_alg.cpp:570-612 (many at once, one consistent state)
void GzAlg::point_queries(const uint32_t* us, const uint32_t* vs, size_t count, uint8_t* outp) {
  for (size_t i = 0; i < count; ++i)
    if (us[i] >= n || vs[i] >= n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
  std::lock_guard<std::mutex> lk(query_mtx);
  ensure_answer();
  if (!count) return;
  qa.ensure(count);
  qb.ensure(count);
  qout.ensure(count);
  GZ_HIP(hipMemcpyAsync(qa.p, us, count * 4, hipMemcpyHostToDevice, st));
  GZ_HIP(hipMemcpyAsync(qb.p, vs, count * 4, hipMemcpyHostToDevice, st));
  launch_point_queries(dsu, qa.p, qb.p, (uint32_t)count, qout.p, st);
  GZ_HIP(hipMemcpyAsync(outp, qout.p, count, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
}

// One consistent state for a group of queries (the serving loop's unit of work): a single read of
// the eager-DSU flag decides between the cached DSU and flush + Borůvka; then every connected-
// components count and point query of the group is answered by one kernel pass and one copy back.
bool GzAlg::answer_queries(const uint32_t* us, const uint32_t* vs, size_t npoint, bool want_cc, uint32_t* ncc_out,
                           uint8_t* point_out) {
  for (size_t i = 0; i < npoint; ++i)
    if (us[i] >= n || vs[i] >= n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
  std::lock_guard<std::mutex> lk(query_mtx);
  // Whether the eager DSU still answers is decided on the device when the host cannot tell: the
  // query is enqueued gated (flush, then a Borůvka that keeps a valid eager DSU), so no host
  // synchronisation precedes it.  Host-side pre_insert calls are applied first.
  bool may_be_cached = false;
  if (cfg.eager_dsu) {
    {
      std::lock_guard<std::mutex> g(eager_mtx);
      if (dsu_valid) submit_eager_locked();
      else eager_used = 0;
    }
    may_be_cached = dsu_valid && eager_live();
  }
  bool cached = false;
  // The pairs go to mapped host memory, where the answer kernel reads them and leaves the error
  // flags and the answers (no copy launches around it).  The previous group's kernel has finished:
  // every group ends with a stream synchronisation.
  aq_in.ensure(std::max<size_t>(1, 2 * npoint));
  aq_out.ensure(kNumFlags + (npoint + 3) / 4);
  if (npoint) {
    std::memcpy(aq_in.p, us, npoint * 4);
    std::memcpy(aq_in.p + npoint, vs, npoint * 4);
  }
  // the answer work: point queries, the component count, the error flags
  const std::function<void()> enqueue_answers = [&] {
    launch_answer_points(dsu, aq_in.d, (uint32_t)npoint, aq_out.d, st);
    if (want_cc) {
      launch_labels(dsu, labels.p, labels.p + n, st);
      GZ_HIP(hipMemcpyAsync(hsmall.p, labels.p + n, 4, hipMemcpyDeviceToHost, st));
    }
  };
  bool answered = false;
  {
    std::lock_guard<std::mutex> g(stage_mtx);
    if (cfg.batched) submit_stage_locked();
    flush_pending();
  }
  // The answers are enqueued with the first group of rounds; they stand when that group finished
  // the query (the usual case), else they are enqueued again after the last round.
  bool skipped = false;
  answered = boruvka(&enqueue_answers, may_be_cached, &skipped);
  cached = skipped;
  if (!answered) {
    enqueue_answers();
    GZ_HIP(hipStreamSynchronize(st));
  }
  if (want_cc && ncc_out) *ncc_out = hsmall.p[0];
  if (npoint) std::memcpy(point_out, reinterpret_cast<const uint8_t*>(aq_out.p + kNumFlags), npoint);
  if (aq_out.p[kFlagBadInput]) {
    launch_fill_u32(flags.p + kFlagBadInput, 0, 1, st);
    GZ_HIP(hipStreamSynchronize(st));
    throw GzError(GZ_ERR_INVALID_ARG, "ingested device pairs named a vertex >= num_vertices; those pairs were skipped");
  }
  return cached;
}

// calc_spanning_forest, src/cc_sketch_alg.cpp:559-568: every root linked under another records the
// forest edge that linked it; the forest is the list of those edges.
size_t GzAlg::spanning_forest(uint32_t* edges, size_t cap) {
  connected_components(nullptr);
  std::vector<uint64_t> slots(n);
  GZ_HIP(hipMemcpyAsync(slots.data(), sf_edge.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
  size_t k = 0;
  for (uint32_t v = 0; v < n; ++v) {
    if (slots[v] == ~0ull) continue;
    if (edges && k < cap) {
      edges[2 * k] = (uint32_t)(slots[v] >> 32);
      edges[2 * k + 1] = (uint32_t)slots[v];
    }
    ++k;
  }
  return k;
}

void GzAlg::dump(uint32_t v0, uint32_t count, void* outp) {
  if ((uint64_t)v0 + count > n) throw GzError(GZ_ERR_INVALID_ARG, "vertex range out of bounds");
  flush_all();
  const size_t sb = sketch_bytes();
  const uint32_t chunk = (uint32_t)std::max<size_t>(1, (size_t(64) << 20) / sb);
  DevBuf<uint8_t> tmp;
  tmp.alloc((size_t)std::min(chunk, count ? count : 1) * sb);
  for (uint32_t i = 0; i < count; i += chunk) {
    const uint32_t c = std::min(chunk, count - i);
    launch_pack(gb, slab.p, geo, v0 + i, c, tmp.p, st);
    GZ_HIP(hipMemcpyAsync((uint8_t*)outp + (size_t)i * sb, tmp.p, (size_t)c * sb, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
  }
}

uint64_t GzAlg::digest() {
  flush_all();
  DevBuf<uint64_t> d;
  d.alloc(1);
  launch_digest(gb, slab.p, geo, d.p, st);
  uint64_t h = 0;
  GZ_HIP(hipMemcpyAsync(&h, d.p, 8, hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
  return h;
}

void GzAlg::load_packed(uint32_t v0, uint32_t count, const void* in) {
  const size_t sb = sketch_bytes();
  const uint32_t chunk = (uint32_t)std::max<size_t>(1, (size_t(64) << 20) / sb);
  DevBuf<uint8_t> tmp;
  tmp.alloc((size_t)std::min(chunk, count ? count : 1) * sb);
  for (uint32_t i = 0; i < count; i += chunk) {
    const uint32_t c = std::min(chunk, count - i);
    GZ_HIP(hipMemcpyAsync(tmp.p, (const uint8_t*)in + (size_t)i * sb, (size_t)c * sb, hipMemcpyHostToDevice, st));
    launch_unpack(gb, slab.p, geo, v0 + i, c, tmp.p, st);
    GZ_HIP(hipStreamSynchronize(st));
  }
  zero_sum = 0;
}

}  // namespace gz

// ==================================================================================================
// C ABI
// ==================================================================================================
using gz::GzAlg;
using gz::GzError;
using gz::GzSketch;

static thread_local std::string g_last_error;

int gz::guarded_call(const std::function<void()>& f) {
  try {
    f();
    g_last_error.clear();
    return GZ_OK;
  } catch (const GzError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return GZ_ERR_OOM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GZ_ERR_DEVICE;
  }
}

int gz::status_of(const std::exception& e) {
  g_last_error = e.what();
  if (const auto* ge = dynamic_cast<const GzError*>(&e)) return ge->code;
  if (dynamic_cast<const std::bad_alloc*>(&e)) return GZ_ERR_OOM;
  return GZ_ERR_DEVICE;
}
void gz::clear_last_error() { g_last_error.clear(); }

template <class F>
static int guarded(F&& f) {
  return gz::guarded_call(std::function<void()>(std::forward<F>(f)));
}

#define GZ_REQUIRE(cond, msg)                                  \
  do {                                                         \
    if (!(cond)) throw GzError(GZ_ERR_INVALID_ARG, msg);       \
  } while (0)

extern "C" {

const char* gz_last_error(void) { return g_last_error.c_str(); }

int gz_device_count(int* out) {
  return guarded([&] {
    GZ_REQUIRE(out, "null out");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *out = c;
  });
}

uint64_t gz_calc_vector_length(uint32_t n) { return gz::calc_vector_length(n); }
size_t gz_calc_cc_samples(uint32_t n, double f) { return gz::calc_cc_samples(n, f); }
size_t gz_calc_bkt_per_col(uint64_t len) { return gz::calc_bkt_per_col(len); }

// ---- sketch ----
int gz_sketch_create(uint64_t vector_len, uint64_t seed, size_t num_samples, size_t cols, int gamma_bytes,
                     gz_sketch** out) {
  return guarded([&] {
    GZ_REQUIRE(out, "null out");
    *out = nullptr;
    auto* s = new GzSketch(vector_len, seed, num_samples, cols, gamma_bytes);
    *out = new gz_sketch{s};
  });
}
int gz_sketch_destroy(gz_sketch* s) {
  return guarded([&] {
    if (!s) return;
    delete s->s;
    delete s;
  });
}
int gz_sketch_update(gz_sketch* s, uint64_t idx) {
  return guarded([&] { GZ_REQUIRE(s, "null sketch"); s->s->update_batch(&idx, 1); });
}
int gz_sketch_update_batch(gz_sketch* s, const uint64_t* idx, size_t count) {
  return guarded([&] {
    GZ_REQUIRE(s && (idx || !count), "null argument");
    s->s->update_batch(idx, count);
  });
}
int gz_sketch_sample(gz_sketch* s, uint64_t* idx, int* result) {
  return guarded([&] {
    GZ_REQUIRE(s && idx && result, "null argument");
    *result = s->s->sample(idx);
  });
}
int gz_sketch_merge(gz_sketch* dst, const gz_sketch* src) {
  return guarded([&] { GZ_REQUIRE(dst && src, "null sketch"); dst->s->merge(*src->s); });
}
int gz_sketch_range_merge(gz_sketch* dst, const gz_sketch* src, size_t start, size_t ns) {
  return guarded([&] { GZ_REQUIRE(dst && src, "null sketch"); dst->s->range_merge(*src->s, start, ns); });
}
int gz_sketch_merge_raw_bucket_buffer(gz_sketch* s, const void* raw) {
  return guarded([&] { GZ_REQUIRE(s && raw, "null argument"); s->s->merge_raw(raw); });
}
int gz_sketch_zero_contents(gz_sketch* s) {
  return guarded([&] { GZ_REQUIRE(s, "null sketch"); s->s->zero(); });
}
int gz_sketch_reset_sample_state(gz_sketch* s) {
  return guarded([&] { GZ_REQUIRE(s, "null sketch"); s->s->sample_idx = 0; });
}
int gz_sketch_serialize(const gz_sketch* s, void* out, size_t bytes) {
  return guarded([&] { GZ_REQUIRE(s && out, "null argument"); s->s->serialize(out, bytes); });
}
int gz_sketch_deserialize(gz_sketch* s, const void* in, size_t bytes) {
  return guarded([&] { GZ_REQUIRE(s && in, "null argument"); s->s->deserialize(in, bytes); });
}
int gz_sketch_equal(const gz_sketch* a, const gz_sketch* b, int* eq) {
  return guarded([&] {
    GZ_REQUIRE(a && b && eq, "null argument");
    if (a->s->geo.nb != b->s->geo.nb || a->s->geo.seed != b->s->geo.seed || a->s->gb != b->s->gb) {
      *eq = 0;
      return;
    }
    std::vector<uint8_t> x(a->s->bytes()), y(b->s->bytes());
    a->s->serialize(x.data(), x.size());
    b->s->serialize(y.data(), y.size());
    *eq = x == y ? 1 : 0;
  });
}
size_t gz_sketch_bucket_array_bytes(const gz_sketch* s) { return s ? s->s->bytes() : 0; }
size_t gz_sketch_get_buckets(const gz_sketch* s) { return s ? s->s->geo.nb : 0; }
size_t gz_sketch_get_columns(const gz_sketch* s) { return s ? s->s->geo.C : 0; }
size_t gz_sketch_get_num_samples(const gz_sketch* s) { return s ? s->s->geo.samples : 0; }
size_t gz_sketch_get_bkt_per_col(const gz_sketch* s) { return s ? s->s->geo.D : 0; }
uint64_t gz_sketch_get_seed(const gz_sketch* s) { return s ? s->s->geo.seed : 0; }

// ---- algorithm ----
void gz_alg_config_default(gz_alg_config* c) {
  if (!c) return;
  c->sketches_factor = 1;
  c->batch_factor = 1;
  c->gamma_bytes = 4;
  c->eager_dsu = 1;
  c->batched = 1;
  c->tile = 0;
  c->flush_threshold = 0;
}
int gz_alg_create(uint32_t n, uint64_t seed, const gz_alg_config* cfg, gz_alg** out) {
  return guarded([&] {
    GZ_REQUIRE(out, "null out");
    *out = nullptr;
    gz_alg_config c;
    gz_alg_config_default(&c);
    if (cfg) c = *cfg;
    auto* a = new GzAlg(n, seed, c);
    *out = new gz_alg{a};
  });
}
int gz_alg_destroy(gz_alg* a) {
  return guarded([&] {
    if (!a) return;
    delete a->a;
    delete a;
  });
}
int gz_alg_allocate_worker_memory(gz_alg* a, size_t) {
  // Worker scratch is the device staging, sized at construction; nothing per worker is needed.
  return guarded([&] { GZ_REQUIRE(a, "null alg"); });
}
size_t gz_alg_get_desired_updates_per_batch(const gz_alg* a) { return a ? a->a->desired_updates_per_batch() : 0; }
uint32_t gz_alg_get_num_vertices(const gz_alg* a) { return a ? a->a->n : 0; }
uint64_t gz_alg_get_seed(const gz_alg* a) { return a ? a->a->seed : 0; }
size_t gz_alg_max_rounds(const gz_alg* a) { return a ? a->a->geo.samples : 0; }
size_t gz_alg_sketch_bytes(const gz_alg* a) { return a ? a->a->sketch_bytes() : 0; }
int gz_alg_pre_insert(gz_alg* a, uint32_t src, uint32_t dst, int, int) {
  return guarded([&] { GZ_REQUIRE(a, "null alg"); a->a->pre_insert(src, dst); });
}
int gz_alg_apply_update_batch(gz_alg* a, int, uint32_t src, const uint32_t* dsts, size_t count) {
  return guarded([&] {
    GZ_REQUIRE(a && (dsts || !count), "null argument");
    for (size_t i = 0; i < count; ++i)
      if (dsts[i] >= a->a->n) throw GzError(GZ_ERR_INVALID_ARG, "destination vertex out of range");
    a->a->apply_update_batch(src, dsts, count);
  });
}
int gz_alg_apply_raw_buckets_update(gz_alg* a, uint32_t src, const void* raw) {
  return guarded([&] { GZ_REQUIRE(a && raw, "null argument"); a->a->apply_raw(src, raw); });
}
int gz_alg_update(gz_alg* a, uint32_t src, uint32_t dst, int) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    GZ_REQUIRE(src < a->a->n && dst < a->a->n, "vertex id out of range");
    a->a->pre_insert(src, dst);
    a->a->apply_update_batch(src, &dst, 1);
    a->a->apply_update_batch(dst, &src, 1);
  });
}
int gz_alg_ingest_pairs_device(gz_alg* a, const uint32_t* pairs, size_t count) {
  return guarded([&] {
    GZ_REQUIRE(a && (pairs || !count), "null argument");
    a->a->ingest_device(pairs, count);
  });
}
int gz_alg_ingest_pairs(gz_alg* a, const uint32_t* pairs, size_t count) {
  return guarded([&] {
    GZ_REQUIRE(a && (pairs || !count), "null argument");
    const uint32_t n = a->a->n;
    for (size_t i = 0; i < 2 * count; ++i)
      if (pairs[i] >= n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
    a->a->ingest_host(pairs, count);
  });
}
int gz_alg_has_cached_query(gz_alg* a, int code, int* out) {
  return guarded([&] {
    GZ_REQUIRE(a && out, "null argument");
    *out = a->a->has_cached_query(code) ? 1 : 0;
  });
}
int gz_alg_prep_query(gz_alg* a, int code) {
  return guarded([&] { GZ_REQUIRE(a, "null alg"); a->a->prep_query(code); });
}
int gz_alg_connected_components(gz_alg* a, uint32_t* parent_out, uint32_t* num_cc) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    uint32_t c = a->a->connected_components(parent_out);
    if (num_cc) *num_cc = c;
  });
}
int gz_alg_point_query(gz_alg* a, uint32_t u, uint32_t v, int* connected) {
  return guarded([&] {
    GZ_REQUIRE(a && connected, "null argument");
    uint8_t r = 0;
    a->a->point_queries(&u, &v, 1, &r);
    *connected = r;
  });
}
int gz_alg_point_queries(gz_alg* a, const uint32_t* us, const uint32_t* vs, size_t count, uint8_t* connected) {
  return guarded([&] {
    GZ_REQUIRE(a && ((us && vs && connected) || !count), "null argument");
    a->a->point_queries(us, vs, count, connected);
  });
}
int gz_alg_calc_spanning_forest(gz_alg* a, uint32_t* edges, size_t cap, size_t* num_edges) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    size_t k = a->a->spanning_forest(edges, cap);
    if (num_edges) *num_edges = k;
  });
}
int gz_alg_write_binary(gz_alg* a, const char* path) {
  return guarded([&] {
    GZ_REQUIRE(a && path, "null argument");
    GzAlg& A = *a->a;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw GzError(GZ_ERR_IO, std::string("cannot open ") + path);
    const uint64_t seed = A.seed;
    const uint32_t nv = A.n;
    const double sf = A.cfg.sketches_factor;
    f.write((const char*)&seed, 8);
    f.write((const char*)&nv, 4);
    f.write((const char*)&sf, 8);
    const size_t sb = A.sketch_bytes();
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, (size_t(64) << 20) / sb);
    std::vector<uint8_t> buf((size_t)std::min(chunk, nv) * sb);
    for (uint32_t v = 0; v < nv; v += chunk) {
      const uint32_t c = std::min(chunk, nv - v);
      A.dump(v, c, buf.data());
      f.write((const char*)buf.data(), (std::streamsize)((size_t)c * sb));
    }
    if (!f) throw GzError(GZ_ERR_IO, std::string("write failed: ") + path);
  });
}
int gz_alg_construct_from_serialized_data(const char* path, const gz_alg_config* cfg, gz_alg** out) {
  return guarded([&] {
    GZ_REQUIRE(path && out, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) throw GzError(GZ_ERR_IO, std::string("cannot open ") + path);
    uint64_t seed = 0;
    uint32_t nv = 0;
    double sf = 1;
    f.read((char*)&seed, 8);
    f.read((char*)&nv, 4);
    f.read((char*)&sf, 8);
    if (!f) throw GzError(GZ_ERR_IO, "truncated header");
    gz_alg_config c;
    gz_alg_config_default(&c);
    if (cfg) c = *cfg;
    c.sketches_factor = sf;  // cc_sketch_alg.cpp:41
    auto* A = new GzAlg(nv, seed, c);
    try {
      const size_t sb = A->sketch_bytes();
      const uint32_t chunk = (uint32_t)std::max<size_t>(1, (size_t(64) << 20) / sb);
      std::vector<uint8_t> buf((size_t)std::min(chunk, nv) * sb);
      for (uint32_t v = 0; v < nv; v += chunk) {
        const uint32_t cnt = std::min(chunk, nv - v);
        f.read((char*)buf.data(), (std::streamsize)((size_t)cnt * sb));
        if (!f) throw GzError(GZ_ERR_IO, "truncated sketch data");
        A->load_packed(v, cnt, buf.data());
      }
      A->dsu_valid = false;  // cc_sketch_alg.cpp:62-63
    } catch (...) {
      delete A;
      throw;
    }
    *out = new gz_alg{A};
  });
}
int gz_alg_dump_sketches(gz_alg* a, uint32_t v0, uint32_t count, void* out, size_t bytes) {
  return guarded([&] {
    GZ_REQUIRE(a && out, "null argument");
    GZ_REQUIRE(bytes >= (size_t)count * a->a->sketch_bytes(), "output buffer too small");
    a->a->dump(v0, count, out);
  });
}
int gz_alg_set_profiling(gz_alg* a, int enable) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    GzAlg& A = *a->a;
    A.flush_all();
    if (!A.prof_bytes.p) {
      A.prof_bytes.alloc(gz::kProfSlots);
      gz::launch_fill_u64(A.prof_bytes.p, 0, gz::kProfSlots, A.st);
      GZ_HIP(hipStreamSynchronize(A.st));
    }
    A.prof = enable != 0;
    A.profile = gz_alg_profile{};
  });
}
size_t gz_alg_last_query_round_sizes(gz_alg* a, uint32_t* entries, uint32_t* giant_root, size_t cap) {
  size_t k = 0;
  const int rc = guarded([&] {
    GZ_REQUIRE(a, "null alg");
    GzAlg& A = *a->a;
    std::lock_guard<std::mutex> lk(A.query_mtx);
    uint32_t h[128];
    GZ_HIP(hipMemcpyAsync(h, A.r_stats.p, sizeof(h), hipMemcpyDeviceToHost, A.st));
    GZ_HIP(hipStreamSynchronize(A.st));
    const size_t rounds = std::min<size_t>(A.last_rounds, 63);
    for (size_t r = 1; r <= rounds && k < cap; ++r, ++k) {
      if (entries) entries[k] = h[2 * r];
      if (giant_root) giant_root[k] = h[2 * r + 1];
    }
  });
  return rc == 0 ? k : 0;
}
int gz_alg_set_query_graphs(gz_alg* a, int enable) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    a->a->use_graphs = enable != 0;
  });
}
int gz_autotune(uint32_t n, uint64_t seed, const gz_alg_config* base, const uint32_t* pairs_device, size_t count,
                gz_alg_config* out, gz_autotune_report* report) {
  return guarded([&] {
    GZ_REQUIRE(out && pairs_device && count, "null argument or empty sample");
    gz_alg_config b;
    gz_alg_config_default(&b);
    if (base) b = *base;
    b.eager_dsu = 0;  // time the sketch update path only
    b.batched = 1;
    // candidates: tiles of 256..2048 destinations; flush thresholds of a quarter, half and all of
    // the sample (at least 2^16 updates), so the sample is cut into 4, 2 or 1 launches
    const uint32_t tiles[] = {256, 512, 1024, 2048};
    std::vector<uint64_t> flushes;
    for (uint64_t div : {4ull, 2ull, 1ull}) {
      const uint64_t f = std::max<uint64_t>(uint64_t(1) << 16, ((uint64_t)count + div - 1) / div);
      if (std::find(flushes.begin(), flushes.end(), f) == flushes.end()) flushes.push_back(f);
    }
    gz_autotune_report rep{};
    double best = -1.0;
    gz_alg_config pick = b;
    for (uint64_t f : flushes)
      for (uint32_t t : tiles) {
        if (rep.trials >= GZ_AUTOTUNE_MAX_TRIALS) break;
        gz_alg_config c = b;
        c.tile = t;
        c.flush_threshold = f;
        GzAlg scratch(n, seed, c);
        double ms = -1.0;
        for (int rep_i = 0; rep_i < 3; ++rep_i) {  // a warm-up ingest, then two timed ones
          const auto t0 = std::chrono::steady_clock::now();
          scratch.ingest_device(pairs_device, count);
          scratch.flush_all();  // synchronises
          const double d = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
          if (rep_i > 0 && (ms < 0 || d < ms)) ms = d;
        }
        rep.tile[rep.trials] = t;
        rep.flush_threshold[rep.trials] = f;
        rep.ms[rep.trials] = ms;
        ++rep.trials;
        if (best < 0 || ms < best) {
          best = ms;
          pick = c;
        }
      }
    if (base) {  // keep the caller's eager-DSU and form choices
      pick.eager_dsu = base->eager_dsu;
      pick.batched = base->batched;
    } else {
      gz_alg_config d;
      gz_alg_config_default(&d);
      pick.eager_dsu = d.eager_dsu;
      pick.batched = d.batched;
    }
    *out = pick;
    if (report) *report = rep;
  });
}
int gz_alg_set_round_finish(gz_alg* a, int enable) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    a->a->use_finish = enable != 0;
  });
}
int gz_alg_set_deep_skip(gz_alg* a, int enable) {
  return guarded([&] {
    GZ_REQUIRE(a, "null alg");
    a->a->use_deep_skip = enable != 0;
  });
}
uint32_t gz_alg_last_query_finish_round(const gz_alg* a) { return a ? a->a->last_finish_start : 0; }
int gz_alg_get_profile(gz_alg* a, gz_alg_profile* out) {
  return guarded([&] {
    GZ_REQUIRE(a && out, "null argument");
    GzAlg& A = *a->a;
    GZ_HIP(hipStreamSynchronize(A.st));
    if (A.prof) A.prof_collect();
    *out = A.profile;
  });
}
int gz_alg_digest(gz_alg* a, uint64_t* out) {
  return guarded([&] {
    GZ_REQUIRE(a && out, "null argument");
    *out = a->a->digest();
  });
}
size_t gz_alg_last_query_rounds(const gz_alg* a) { return a ? a->a->last_rounds : 0; }
uint64_t gz_alg_get_total_updates(const gz_alg* a) { return a ? a->a->total_updates.load() : 0; }
uint64_t gz_alg_footprint_bytes(uint32_t n, const gz_alg_config* cfg) {
  gz_alg_config c;
  gz_alg_config_default(&c);
  if (cfg) c = *cfg;
  if (n < 2) return 0;
  return GzAlg::footprint(n, c);
}

}  // extern "C"
