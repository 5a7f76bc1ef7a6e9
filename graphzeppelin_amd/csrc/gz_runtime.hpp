// gz_runtime.hpp -- host-side runtime objects behind the C ABI: device sketch slabs, staging,
// the eager DSU and the Borůvka query loop.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gz_hip.h"
#include "gz_kernels.hpp"

namespace gz {

struct GzError : std::runtime_error {
  int code;
  GzError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define GZ_HIP(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess)                                                                           \
      throw ::gz::GzError(e_ == hipErrorOutOfMemory ? GZ_ERR_OOM : GZ_ERR_DEVICE,                   \
                          std::string(#x) + " failed: " + hipGetErrorString(e_));                   \
  } while (0)

// Owning device buffer.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    GZ_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Owning pinned host buffer.
template <class T>
struct PinBuf {
  T* p = nullptr;
  size_t n = 0;
  PinBuf() = default;
  PinBuf(const PinBuf&) = delete;
  PinBuf& operator=(const PinBuf&) = delete;
  PinBuf(PinBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  ~PinBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    GZ_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

// Owning pinned host buffer that kernels read and write in place (fine-grained, mapped): for a
// few words per query, a kernel access over the bus replaces a copy launch.
template <class T>
struct MapBuf {
  T* p = nullptr;  // host address
  T* d = nullptr;  // device address
  size_t n = 0;
  MapBuf() = default;
  MapBuf(const MapBuf&) = delete;
  MapBuf& operator=(const MapBuf&) = delete;
  ~MapBuf() { release(); }
  void ensure(size_t count) {
    if (count <= n) return;
    release();
    GZ_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
    GZ_HIP(hipHostGetDevicePointer((void**)&d, p, 0));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = d = nullptr;
    n = 0;
  }
};

int device_cu_count();
void require_device();

// Standalone Sketch (include/sketch.h:40-191) with its buckets in HBM.
class GzSketch {
 public:
  Geo geo{};
  int gb;
  int dev = 0;  // the device the sketch's buffers and stream live on
  size_t sample_idx = 0;
  hipStream_t st = nullptr;
  DevBuf<uint8_t> slab;  // packed bucket records (see Geo::row)
  DevBuf<uint64_t> col_bf, raw;
  DevBuf<uint32_t> dsts, out;
  DevBuf<Tile> tiles;
  DevBuf<uint8_t> packed;

  GzSketch(uint64_t vector_len, uint64_t seed, size_t samples, size_t cols, int gamma_bytes);
  ~GzSketch();
  void update_batch(const uint64_t* idx, size_t count);
  int sample(uint64_t* idx);
  void merge(const GzSketch& o);
  void range_merge(const GzSketch& o, size_t start, size_t n_samples);
  void merge_raw(const void* raw_buckets);
  void zero();
  void serialize(void* out, size_t bytes) const;
  void deserialize(const void* in, size_t bytes);
  size_t bytes() const { return (size_t)geo.nb * (8 + gb); }
};

// CCSketchAlg (include/cc_sketch_alg.h) with every vertex sketch resident in HBM.
class GzAlg {
 public:
  GzAlg(uint32_t n, uint64_t seed, const gz_alg_config& cfg);
  ~GzAlg();

  uint32_t n;
  uint64_t seed;
  gz_alg_config cfg;
  Geo geo{};
  int gb;
  uint32_t tile;
  uint32_t grid_cap;
  hipStream_t st = nullptr;

  // sketch slab
  DevBuf<uint8_t> slab;  // packed bucket records, one 256-B aligned row per vertex
  DevBuf<uint64_t> col_bf;
  // DSU / forest
  DevBuf<uint32_t> parent, flags, labels, qa, qb;
  DevBuf<uint64_t> sf_edge;
  DevBuf<uint64_t> sf_hash;
  DevBuf<uint8_t> qout;
  DsuDev dsu{};
  // Borůvka round buffers
  DevBuf<uint32_t> r_cnt, r_off, r_cursor, r_members[2], r_mroot[2], r_totals, r_stats, mod;
  DevBuf<RoundSeg> r_segs[2];
  DevBuf<uint2> r_nd;
  DevBuf<uint64_t> r_xacc, r_samp;
  DevBuf<uint8_t> r_det, r_hint;
  DevBuf<uint32_t> r_xcnt;
  RoundBufs rb{};
  // device stream ingest (pending pairs -> grouping -> tiles)
  size_t cap_pairs;
  size_t pending_n = 0;  // pairs copied into `pending`
  const uint32_t* ref_pairs = nullptr;  // pairs referenced in the caller's memory (ingest_device_stable)
  size_t ref_n = 0;
  DevBuf<uint32_t> pending, dst_arr, g_cnt, g_off, g_toff, g_cursor, g_totals, g_hist, g_hoff, g_htotals;
  DevBuf<uint2> g_rec;
  DevBuf<uint64_t> g_hpartials;
  DevBuf<uint64_t> g_partials;
  DevBuf<Tile> g_tiles;
  DevBuf<uint8_t> merged;
  GroupBufs gbufs{};
  // host staging for apply_update_batch (batched form: two buffers, plain form: a ring)
  struct Stage {
    PinBuf<uint8_t> host;  // [tiles | dsts]
    DevBuf<uint8_t> dev;
    hipEvent_t ev = nullptr;
    bool inflight = false;
    std::vector<uint32_t> ent_v, ent_off, ent_cnt;
    size_t used = 0;       // destinations staged
    size_t dcap = 0;       // destination capacity
    size_t max_tiles = 0;  // tile capacity
  };
  static constexpr size_t kMaxEntries = 65536;
  std::mutex stage_mtx;
  std::vector<Stage> stages;
  size_t cur_stage = 0;
  size_t stage_cap;
  std::vector<uint32_t> vstamp, vtiles;
  uint32_t epoch = 0;
  // host ingest: copy stream, pinned staging (two buffers of kHostStagePairs pairs)
  static constexpr size_t kHostStagePairs = size_t(1) << 20;
  hipStream_t cst = nullptr;
  struct HostStage {
    PinBuf<uint32_t> buf;
    hipEvent_t ev = nullptr;  // recorded on the copy stream after the buffer's last copy
    bool busy = false;
  };
  HostStage hstage[2];
  int hstage_cur = 0;
  hipEvent_t copy_ev = nullptr;  // the latest host copy; the compute stream waits on it
  hipEvent_t pend_free_ev = nullptr;  // a flush's grouping has read `pending` (host copies may overwrite it)
  bool pend_wait = false;             // the next host copy into `pending` must wait for pend_free_ev
  void ensure_host_ingest();
  // eager staging (pre_insert)
  std::mutex eager_mtx;
  PinBuf<uint32_t> eager_host;
  DevBuf<uint32_t> eager_dev;
  DevBuf<uint8_t> eager_merged;
  hipEvent_t eager_ev = nullptr;
  size_t eager_used = 0, eager_cap;
  PinBuf<uint32_t> hflags;  // device flags, or a query's round flags
  PinBuf<uint32_t> hsmall;
  MapBuf<uint32_t> eager_inv;  // mapped mirror of kFlagEagerInvalid, cleared after each query
  MapBuf<uint32_t> qforce;     // DsuDev::query_force: 1 runs the next query whatever the eager DSU's state
  bool eager_live() const { return !(eager_inv.p && *reinterpret_cast<const volatile uint32_t*>(eager_inv.p)); }
  MapBuf<uint32_t> aq_in;  // a query group's point queries: us then vs, read by the answer kernel
  MapBuf<uint32_t> aq_out;  // the error flags, then one answer byte per point query, written by it

  // kernel timing (gz_alg_set_profiling)
  struct Timed {
    hipEvent_t a, b;
    int kind;          // 0 update kernel, 1 Borůvka round
    double host_bytes;  // algorithmic bytes known on the host (staged paths / rounds)
    bool dev_bytes;     // bytes accumulated on the device (grouped flush)
    uint64_t items;
  };
  bool prof = false;
  std::vector<Timed> timed;
  std::vector<hipEvent_t> ev_pool;
  DevBuf<uint64_t> prof_bytes;  // device accumulator for grouped-flush bytes
  gz_alg_profile profile{};
  hipEvent_t prof_begin(int kind, double host_bytes, bool dev_bytes, uint64_t items);
  void prof_end();
  void prof_collect();
  hipEvent_t take_event();

  std::atomic<bool> dsu_valid{true};
  std::atomic<bool> update_locked{false};
  size_t last_rounds = 0;
  // Whether the vertex sketches XOR to zero (true for sketches built from edge updates only):
  // 1 known zero, 0 unknown (checked at the next query), -1 known not zero.
  std::atomic<int> zero_sum{1};
  bool sketches_sum_to_zero();
  std::atomic<uint64_t> total_updates{0};
  std::mutex query_mtx;

  size_t sketch_bytes() const { return (size_t)geo.nb * (8 + gb); }
  size_t desired_updates_per_batch() const {
    size_t num = sketch_bytes() / sizeof(uint32_t);
    return (size_t)(num * cfg.batch_factor);
  }
  static uint64_t footprint(uint32_t n, const gz_alg_config& cfg);
  double batch_alg_bytes(const std::vector<uint32_t>& counts) const;

  void pre_insert(uint32_t src, uint32_t dst);
  void apply_update_batch(uint32_t src, const uint32_t* dsts, size_t k);
  void apply_raw(uint32_t src, const void* raw);
  void ingest_device(const uint32_t* pairs_dev, size_t count);
  // ingest_device for a caller whose pairs stay valid and unchanged until release_stable(): the
  // flush groups them where they are (consecutive calls on consecutive ranges of one buffer form
  // one range), and the eager DSU reads them in place, so nothing is copied per call.
  void ingest_device_stable(const uint32_t* pairs_dev, size_t count);
  // Applies (or, if that fails, forgets) every referenced range: after it returns nothing reads
  // memory given to ingest_device_stable.
  void release_stable();
  // (src,dst) pairs in host memory.  Each chunk is DMA'd into the pending buffer on a copy stream
  // that overlaps the kernels: straight from the caller's memory when it is pinned, else through
  // two pinned staging buffers.  stable: the caller keeps the pairs valid and unchanged until the
  // next flush, so the call need not wait for its DMA; otherwise a call whose pairs were read in
  // place returns only after their copy has landed.
  void ingest_host(const uint32_t* pairs, size_t count, bool stable = false);
  bool has_cached_query(int code);
  void prep_query(int code);
  uint32_t connected_components(uint32_t* parent_out);
  void point_queries(const uint32_t* us, const uint32_t* vs, size_t count, uint8_t* out);
  // Answers a group of queries against one state; returns whether the cached DSU answered them.
  bool answer_queries(const uint32_t* us, const uint32_t* vs, size_t npoint, bool want_cc, uint32_t* ncc_out,
                      uint8_t* point_out);
  size_t spanning_forest(uint32_t* edges, size_t cap);
  void dump(uint32_t v0, uint32_t count, void* out);
  void load_packed(uint32_t v0, uint32_t count, const void* in);
  uint64_t digest();
  void flush_all();

 private:
  void submit_stage_locked();
  void plain_launch(uint32_t src, const uint32_t* dsts, size_t k);
  void flush_pending();
  void submit_eager_locked();
  void sync_eager_state();
  void ensure_answer();
  // Runs Borůvka.  `speculate`, when given, enqueues the caller's answer work (kernels and copies
  // reading the DSU); it runs right after the first group of rounds, and boruvka() returns true when
  // that group finished the query, so the work read the final DSU and one synchronisation served
  // both; false means it has to be enqueued again.
  // gated (with speculate): the device runs the query only if the eager DSU turned invalid;
  // otherwise every kernel of it returns at once, the answers come from the eager DSU, and
  // *skipped is set.
  bool boruvka(const std::function<void()>* speculate = nullptr, bool gated = false, bool* skipped = nullptr);
  uint32_t read_flag(int slot);

  // Query prefix (DSU reset, round 0, rounds 1..K and a single-workgroup finish of the remaining
  // rounds) replayed from a HIP graph: every parameter is fixed per instance (round state lives on
  // the device), so one graph launch replaces ~5 kernel launches per round.  K is the last round
  // the previous query needed launched rounds for: the first round after which a giant existed and
  // at most finish_max() vertices remained outside it (queries on one stream need similar counts),
  // else its round count plus one.  Graphs are built on first use and kept per (K, giant_ok).
 public:
  bool use_graphs = true;
  bool use_finish = true;           // the single-workgroup finish of the last rounds
  bool use_deep_skip = true;        // reduces skip the deep records of members whose deep flag is clear
  uint32_t last_finish_start = 0;   // round at which it took over in the last query (0: none)
  static constexpr uint32_t kFirstGraphRounds = 12;
 private:
  std::map<uint32_t, hipGraphExec_t> query_graphs;
  PinBuf<uint32_t> hround;  // per-round diagnostics of the last query (copied when it ends)
  hipEvent_t hround_ev = nullptr;  // that copy has landed
  bool hround_pending = false;      // hround holds the last query's diagnostics, not yet used
  bool hround_on_event = false;     // ... once hround_ev has completed
  MapBuf<uint32_t> rep;             // the finish kernel's report: round flags, then 128 diagnostic words
  uint32_t next_prefix = 0;  // K of the next query's prefix (0: none chosen yet)
  // The spanning-forest membership set that the eager DSU's deletion check reads is built from the
  // forest edges after a query, but only before the first eager batch that needs it.
  bool sf_pending = false;
  // src != pairs: the batch is copied from src into pairs (device memory) by the same launch
  void eager_batch(const uint32_t* src, uint32_t* pairs, size_t count, uint8_t* merged);
  void enqueue_query_prefix(bool giant_ok, uint32_t last);
  hipGraphExec_t get_query_graph(bool giant_ok, uint32_t last);
};

// Runs f, converting exceptions to a gz_status and the thread's last-error message.
int guarded_call(const std::function<void()>& f);
// The same conversion for entry points that catch their own exceptions: the status of e (set as
// the thread's last-error message), and clear_last_error() after a successful call.
int status_of(const std::exception& e);
void clear_last_error();

}  // namespace gz

// Opaque handle types of the C ABI.
struct gz_sketch {
  gz::GzSketch* s;
};
struct gz_alg {
  gz::GzAlg* a;
};
