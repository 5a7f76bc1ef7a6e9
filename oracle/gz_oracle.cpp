// gz_oracle.cpp -- CPU restatement of GraphZeppelin's streaming connected-components path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: the parity tests,
// the smoke check and the CPU-baseline leg of the benchmark load it; the product never does.
//
// Every function cites the reference file:line it restates (paths relative to the reference
// repository root).  Parity pinning of each part:
//   * xxh3_len8            : golden vectors made with the `xxhash` library (tests/golden), plus
//                            the closed form of XXH3_len_4to8_64b / XXH3_rrmxmx.
//   * geometry             : integer table for n = 1024 .. 2^20 asserted in the tests.
//   * Sketch::update/sample: linearity, insert+delete -> all zero, "index 0 updated 3x -> GOOD 0"
//                            (test/sketch_test.cpp:66-80), raw-buffer merge twice -> ZERO
//                            (test/sketch_test.cpp:423-458), depth histogram ~ 2^-(d+1),
//                            single-update bucket image built from library hashes.
//   * Borůvka / CC         : 78 components on test/res/multiples_graph_1024.txt and on the final
//                            state of test/res/multiples_graph_1024_stream.data; exact DSU over the
//                            final edge-parity set on seeded random streams.
//   * Eager DSU            : the state sequence of test/cc_alg_test.cpp:223-263.
// The byte image of a whole sketch has no golden vector from the reference itself (its core needs
// headers of un-vendored dependencies to compile), so it is pinned by the checks above only.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <omp.h>

namespace oracle {

// ---------------------------------------------------------------------------------------------
// XXH3-64 of exactly 8 input bytes, restated from the public XXH3 specification
// (XXH3_len_4to8_64b followed by XXH3_rrmxmx).  The reference calls XXH3_64bits_withSeed on the
// 8 little-endian bytes of a vec_t (include/types.h:8-9, include/bucket.h:56,62).
// ---------------------------------------------------------------------------------------------
static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static inline uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

static constexpr uint64_t kMx2 = 0x9FB21C651E98DF25ULL;
// bytes [8,16) and [16,24) of the default 192-byte XXH3 secret, read little-endian
static constexpr uint64_t kSecret8 = 0x1cad21f72c81017cULL;
static constexpr uint64_t kSecret16 = 0xdb979083e96dd4deULL;

uint64_t xxh3_len8(uint64_t input, uint64_t seed) {
  seed ^= (uint64_t)bswap32((uint32_t)seed) << 32;
  const uint32_t in_lo = (uint32_t)input;          // readLE32(input)
  const uint32_t in_hi = (uint32_t)(input >> 32);  // readLE32(input + 4)
  const uint64_t bitflip = (kSecret8 ^ kSecret16) - seed;
  const uint64_t input64 = (uint64_t)in_hi + ((uint64_t)in_lo << 32);
  uint64_t h = input64 ^ bitflip;
  // rrmxmx with len = 8
  h ^= rotl64(h, 49) ^ rotl64(h, 24);
  h *= kMx2;
  h ^= (h >> 35) + 8;
  h *= kMx2;
  h ^= h >> 28;
  return h;
}

// ---------------------------------------------------------------------------------------------
// Geometry (include/sketch.h:62-76, 181, 187-190; src/sketch.cpp:8-21), computed with doubles
// exactly as the reference does.
// ---------------------------------------------------------------------------------------------
static const double kNumSamplesDiv = 1 - std::log2(2 - 0.8);  // CameoSketch branch, sketch.h:188

uint64_t calc_vector_length(uint32_t n) {  // sketch.h:62-64
  return (uint64_t)std::ceil(double(n) * (n - 1) / 2);
}
size_t calc_cc_samples(uint32_t n, double f) {  // sketch.h:74-76
  return std::max(size_t(18), (size_t)std::ceil(f * std::log2(n) / kNumSamplesDiv));
}
size_t calc_bkt_per_col(uint64_t len) { return (size_t)(std::ceil(std::log2((double)len)) + 1); }  // sketch.h:181

// concat pairing (src/util.cpp:45-57)
static inline uint64_t concat_pairing(uint32_t i, uint32_t j) {
  if (i > j) std::swap(i, j);
  return ((uint64_t)i << 32) | j;
}

enum SampleResult { GOOD = 0, ZERO = 1, FAIL = 2 };  // include/sketch.h:20-24
struct SketchSample { uint64_t idx; int result; };

struct OutOfSamples : std::runtime_error {
  OutOfSamples() : std::runtime_error("out of samples") {}
};

// Bucket: packed {u64 alpha; G gamma;} (include/bucket.h:7-12)
#pragma pack(push, 1)
template <class G> struct Bucket { uint64_t alpha; G gamma; };
#pragma pack(pop)

// ---------------------------------------------------------------------------------------------
// Sketch (CameoSketch, src/sketch.cpp)
// ---------------------------------------------------------------------------------------------
template <class G>
class Sketch {
 public:
  uint64_t seed;
  size_t num_samples, cols_per_sample, num_columns, bkt_per_col, num_buckets;
  size_t sample_idx = 0;
  std::vector<Bucket<G>> b;

  Sketch(uint64_t vector_len, uint64_t seed_, size_t samples = 1, size_t cols = 1)  // sketch.cpp:8-21
      : seed(seed_), num_samples(samples), cols_per_sample(cols) {
    num_columns = num_samples * cols_per_sample;
    bkt_per_col = calc_bkt_per_col(vector_len);
    num_buckets = num_columns * bkt_per_col + 1;
    b.assign(num_buckets, Bucket<G>{0, 0});
  }
  uint64_t column_seed(size_t c) const { return seed + c * 5; }  // sketch.h:175
  uint64_t checksum_seed() const { return seed; }                // sketch.h:176

  // bucket.h:54-59
  static inline size_t index_depth(uint64_t idx, uint64_t col_seed, size_t max_depth) {
    uint64_t h = xxh3_len8(idx, col_seed);
    h |= (1ull << max_depth);
    return (size_t)__builtin_ctzll(h);
  }
  // bucket.h:61-63
  static inline G index_hash(uint64_t idx, uint64_t s) { return (G)xxh3_len8(idx, s); }
  // bucket.h:65-67
  inline bool is_good(const Bucket<G>& bk) const { return bk.gamma == index_hash(bk.alpha, checksum_seed()); }

  // sketch.cpp:69-83 (CameoSketch: one bucket per column)
  void update(uint64_t idx) {
    G chk = index_hash(idx, checksum_seed());
    b[num_buckets - 1].alpha ^= idx;
    b[num_buckets - 1].gamma ^= chk;
    for (size_t c = 0; c < num_columns; ++c) {
      size_t d = index_depth(idx, column_seed(c), bkt_per_col);
      if (d < bkt_per_col) {
        b[c * bkt_per_col + d].alpha ^= idx;
        b[c * bkt_per_col + d].gamma ^= chk;
      }
    }
  }
  void zero_contents() {  // sketch.cpp:86-92
    for (auto& x : b) x = Bucket<G>{0, 0};
    sample_idx = 0;
  }
  SketchSample sample() {  // sketch.cpp:94-116
    if (sample_idx >= num_samples) throw OutOfSamples();
    size_t idx = sample_idx++;
    size_t first_column = idx * cols_per_sample;
    const Bucket<G>& det = b[num_buckets - 1];
    if (det.alpha == 0 && det.gamma == 0) return {0, ZERO};
    if (is_good(det)) return {det.alpha, GOOD};
    for (size_t i = 0; i < cols_per_sample; ++i)
      for (size_t j = 0; j < bkt_per_col; ++j) {
        const Bucket<G>& bk = b[(i + first_column) * bkt_per_col + j];
        if (is_good(bk)) return {bk.alpha, GOOD};
      }
    return {0, FAIL};
  }
  void merge(const Sketch& o) {  // sketch.cpp:149-154
    for (size_t i = 0; i < num_buckets; ++i) { b[i].alpha ^= o.b[i].alpha; b[i].gamma ^= o.b[i].gamma; }
  }
  void range_merge(const Sketch& o, size_t start_sample, size_t n_samples) {  // sketch.cpp:156-179
    if (start_sample + n_samples > num_samples) { sample_idx = num_samples; return; }
    sample_idx = std::max(sample_idx, start_sample);
    b[num_buckets - 1].alpha ^= o.b[num_buckets - 1].alpha;
    b[num_buckets - 1].gamma ^= o.b[num_buckets - 1].gamma;
    size_t s = start_sample * cols_per_sample * bkt_per_col;
    size_t cnt = n_samples * cols_per_sample * bkt_per_col;
    for (size_t i = 0; i < cnt; ++i) { b[s + i].alpha ^= o.b[s + i].alpha; b[s + i].gamma ^= o.b[s + i].gamma; }
  }
  void merge_raw(const Bucket<G>* raw) {  // sketch.cpp:181-186
    for (size_t i = 0; i < num_buckets; ++i) { b[i].alpha ^= raw[i].alpha; b[i].gamma ^= raw[i].gamma; }
  }
  size_t bytes() const { return num_buckets * sizeof(Bucket<G>); }
};

// ---------------------------------------------------------------------------------------------
// DisjointSetUnion_MT (include/dsu.h:101-174).  Priorities come from a fixed-seed mt19937_64
// instead of the wall clock (dsu.h:120-127); only the partition is ever compared.
// ---------------------------------------------------------------------------------------------
class DSU {
 public:
  uint32_t n;
  std::unique_ptr<std::atomic<uint32_t>[]> parent;
  std::vector<uint64_t> prio;
  explicit DSU(uint32_t n_) : n(n_), parent(new std::atomic<uint32_t>[n_]), prio(n_) {
    std::mt19937_64 g(0x5eed);
    for (uint32_t i = 0; i < n; ++i) { parent[i] = i; prio[i] = g(); }
  }
  uint32_t find_root(uint32_t u) {  // dsu.h:145-152 (path halving)
    while (parent[parent[u]] != parent[u]) {
      parent[u] = parent[parent[u]].load();
      u = parent[u];
    }
    return parent[u];
  }
  bool merge(uint32_t a, uint32_t b) {  // dsu.h:155-167
    while ((a = find_root(a)) != (b = find_root(b))) {
      if (prio[a] == prio[b] && a > b) std::swap(a, b);
      if (prio[a] < prio[b]) std::swap(a, b);
      uint32_t expect = b;
      if (parent[b].compare_exchange_weak(expect, a)) return true;
    }
    return false;
  }
  void reset() { for (uint32_t i = 0; i < n; ++i) parent[i] = i; }  // dsu.h:169-173
};

// ---------------------------------------------------------------------------------------------
// CCSketchAlg (include/cc_sketch_alg.h, src/cc_sketch_alg.cpp)
// ---------------------------------------------------------------------------------------------
template <class G>
class CCAlg {
 public:
  uint32_t n;
  uint64_t seed;
  double sketches_factor, batch_factor;
  bool eager;
  std::vector<std::unique_ptr<Sketch<G>>> sk;
  std::vector<std::mutex> vmtx;
  DSU dsu;
  bool dsu_valid = true;
  std::atomic<bool> shared_dsu_valid{true};
  std::vector<std::unordered_set<uint32_t>> sf;
  std::vector<std::mutex> sf_mtx;
  std::vector<std::unique_ptr<Sketch<G>>> delta;
  size_t last_query_rounds = 0;

  CCAlg(uint32_t n_, uint64_t seed_, double sf_ = 1, double bf = 1, bool eager_ = true)  // cc_sketch_alg.cpp:11-28
      : n(n_), seed(seed_), sketches_factor(sf_), batch_factor(bf), eager(eager_), vmtx(n_), dsu(n_), sf(n_), sf_mtx(n_) {
    uint64_t len = calc_vector_length(n);
    size_t samples = calc_cc_samples(n, sketches_factor);
    sk.reserve(n);
    for (uint32_t i = 0; i < n; ++i) sk.emplace_back(new Sketch<G>(len, seed, samples));
  }
  std::unique_ptr<Sketch<G>> new_sketch() const {
    return std::unique_ptr<Sketch<G>>(new Sketch<G>(calc_vector_length(n), seed, calc_cc_samples(n, sketches_factor)));
  }
  void allocate_worker_memory(size_t w) {  // cc_sketch_alg.h:161-169
    delta.clear();
    for (size_t i = 0; i < w; ++i) delta.push_back(new_sketch());
  }
  size_t desired_updates_per_batch() const {  // cc_sketch_alg.h:146-150
    size_t num = sk[0]->bytes() / sizeof(uint32_t);
    return (size_t)(num * batch_factor);
  }
  void pre_insert(uint32_t a, uint32_t b2) {  // cc_sketch_alg.cpp:79-104
    if (!eager) {
      if (dsu_valid) { dsu_valid = false; shared_dsu_valid = false; }
      return;
    }
    if (!dsu_valid) return;
    uint32_t src = std::min(a, b2), dst = std::max(a, b2);
    std::lock_guard<std::mutex> lk(sf_mtx[src]);
    if (dsu.merge(src, dst)) sf[src].insert(dst);
    else if (sf[src].count(dst)) { dsu_valid = false; shared_dsu_valid = false; }
  }
  void apply_update_batch(int thr, uint32_t src, const uint32_t* dsts, size_t k) {  // cc_sketch_alg.cpp:106-118
    Sketch<G>& d = *delta[thr];
    d.zero_contents();
    for (size_t i = 0; i < k; ++i) d.update(concat_pairing(src, dsts[i]));
    std::lock_guard<std::mutex> lk(vmtx[src]);
    sk[src]->merge(d);
  }
  void apply_raw(uint32_t src, const Bucket<G>* raw) {  // cc_sketch_alg.cpp:120-123
    std::lock_guard<std::mutex> lk(vmtx[src]);
    sk[src]->merge_raw(raw);
  }
  void update(uint32_t a, uint32_t b2) {  // cc_sketch_alg.cpp:127-133
    pre_insert(a, b2);
    uint64_t idx = concat_pairing(a, b2);
    sk[a]->update(idx);
    sk[b2]->update(idx);
  }
  // cc_sketch_alg.cpp:137-166
  bool sample_supernode(Sketch<G>& s) {
    SketchSample smp = s.sample();
    if (smp.result == FAIL) return true;
    if (smp.result == GOOD) {
      uint32_t u = (uint32_t)(smp.idx >> 32), v = (uint32_t)smp.idx;
      // An index decoded from a checksum collision may name vertices out of range; the
      // reference would index out of bounds there.  Treated as a failed sample.
      if (u >= n || v >= n) return true;
      if (dsu.merge(u, v)) {
        uint32_t a = std::min(u, v), b2 = std::max(u, v);
        std::lock_guard<std::mutex> lk(sf_mtx[a]);
        sf[a].insert(b2);
        return true;
      }
    }
    return false;
  }
  // boruvka_emulation (cc_sketch_alg.cpp:464-513).  Round 0 samples every vertex sketch
  // (run_round_zero, 223-244); round r >= 1 builds every supernode's column-r + deterministic
  // bucket as the XOR of its members (perform_boruvka_round + range_merge, 246-380) after
  // regrouping the vertices by DSU root (create_merge_instructions, 382-462).  The per-thread
  // partition / GlobalMergeData machinery of the reference only parallelises this.
  void boruvka() {
    dsu.reset();
    for (auto& s : sf) s.clear();
    size_t round = 0;
    std::vector<uint32_t> order(n), roots(n);
    while (true) {
      bool modified = false;
      if (round == 0) {
        int any = 0;
        bool except = false;
#pragma omp parallel for reduction(| : any) schedule(dynamic, 256)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
          try {
            if (sample_supernode(*sk[i])) any = 1;
          } catch (OutOfSamples&) {
#pragma omp critical
            except = true;
          }
        }
        if (except) throw OutOfSamples();
        modified = any;
      } else {
        if (round >= sk[0]->num_samples) throw OutOfSamples();
        // group by root: counting sort by root id (create_merge_instructions)
        std::vector<uint32_t> cnt(n + 1, 0);
        for (uint32_t v = 0; v < n; ++v) { roots[v] = dsu.find_root(v); cnt[roots[v] + 1]++; }
        for (uint32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
        std::vector<uint32_t> cur(cnt.begin(), cnt.end() - 1);
        for (uint32_t v = 0; v < n; ++v) order[cur[roots[v]]++] = v;
        std::vector<uint32_t> seg_roots;
        for (uint32_t r = 0; r < n; ++r) if (cnt[r + 1] > cnt[r]) seg_roots.push_back(r);
        int any = 0;
#pragma omp parallel reduction(| : any)
        {
          auto local = new_sketch();
#pragma omp for schedule(dynamic, 64)
          for (int64_t si = 0; si < (int64_t)seg_roots.size(); ++si) {
            uint32_t r = seg_roots[si];
            local->zero_contents();
            for (uint32_t p = cnt[r]; p < cnt[r + 1]; ++p) local->range_merge(*sk[order[p]], round, 1);
            if (sample_supernode(*local)) any = 1;
          }
        }
        modified = any;
      }
      if (!modified) break;
      ++round;
    }
    last_query_rounds = round;
    dsu_valid = true;
    shared_dsu_valid = true;
  }
  // connected_components (cc_sketch_alg.cpp:515-557) + ConnectedComponents (return_types.cpp:5-19)
  uint32_t connected_components(uint32_t* parent_out) {
    if (!shared_dsu_valid) {
      try {
        boruvka();
      } catch (...) {
        for (auto& s : sk) s->sample_idx = 0;
        throw;
      }
      for (auto& s : sk) s->sample_idx = 0;
    }
    uint32_t cc = 0;
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t r = dsu.find_root(i);
      if (parent_out) parent_out[i] = r;
      if (r == i) ++cc;
    }
    return cc;
  }
  // point_query (cc_sketch_alg.cpp:570-612)
  bool point_query(uint32_t a, uint32_t b2) {
    if (!dsu_valid) {
      try {
        boruvka();
      } catch (...) {
        for (auto& s : sk) s->sample_idx = 0;
        throw;
      }
      for (auto& s : sk) s->sample_idx = 0;
    }
    return dsu.find_root(a) == dsu.find_root(b2);
  }
};

// Exact CC over the final edge-parity set: the lossless check that GraphVerifier performs
// (test/util/graph_verifier.cpp:49-79), with a hash set instead of a dense bit matrix.
uint32_t exact_cc(uint32_t n, const uint32_t* pairs, size_t count, uint32_t* labels_out) {
  std::unordered_map<uint64_t, uint8_t> par;
  par.reserve(count);
  for (size_t i = 0; i < count; ++i) par[concat_pairing(pairs[2 * i], pairs[2 * i + 1])] ^= 1;
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = i;
  auto find = [&](uint32_t u) { while (p[u] != u) { p[u] = p[p[u]]; u = p[u]; } return u; };
  for (auto& kv : par) {
    if (!kv.second) continue;
    uint32_t a = find((uint32_t)(kv.first >> 32)), b = find((uint32_t)kv.first);
    if (a != b) { if (a < b) p[b] = a; else p[a] = b; }
  }
  uint32_t cc = 0;
  std::vector<uint32_t> minlab(n, UINT32_MAX);
  for (uint32_t i = 0; i < n; ++i) { uint32_t r = find(i); minlab[r] = std::min(minlab[r], i); }
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t r = find(i);
    if (r == i) ++cc;
    if (labels_out) labels_out[i] = minlab[r];
  }
  return cc;
}

// ---------------------------------------------------------------------------------------------
// CPU baseline driver: the reference CPU ingest path restated (include/graph_sketch_driver.h:
// 141-209, include/worker_thread_group.h:63-90).  Every update is inserted twice, once per
// direction (graph_sketch_driver.h:171-172), into per-vertex gutters of
// desired_updates_per_batch() destinations; a full gutter is applied by apply_update_batch with
// the thread's delta sketch and the per-vertex merge lock.  The stream threads double as the
// workers (every host thread does both jobs), and prep_query drains all gutters.
// ---------------------------------------------------------------------------------------------
template <class G>
struct CpuDriver {
  CCAlg<G>& alg;
  size_t nthreads, gsize;
  std::vector<std::vector<uint32_t>> gut;
  std::vector<std::mutex> gmtx;
  std::atomic<uint64_t> total{0};
  CpuDriver(CCAlg<G>& a, size_t nt) : alg(a), nthreads(nt), gut(a.n), gmtx(a.n) {
    gsize = std::max<size_t>(1, alg.desired_updates_per_batch());
    alg.allocate_worker_memory(nt);
    for (auto& g : gut) g.reserve(0);
  }
  void insert(int thr, uint32_t src, uint32_t dst, std::vector<uint32_t>& spare) {
    std::vector<uint32_t> full;
    {
      std::lock_guard<std::mutex> lk(gmtx[src]);
      auto& g = gut[src];
      if (g.capacity() < gsize) g.reserve(gsize);
      g.push_back(dst);
      if (g.size() < gsize) return;
      full.swap(g);
      g.swap(spare);
      g.clear();
    }
    total += full.size();
    alg.apply_update_batch(thr, src, full.data(), full.size());
    spare.swap(full);
  }
  void process(const uint32_t* pairs, size_t count) {  // process_stream_until
    auto task = [&](int t) {
      std::vector<uint32_t> spare;
      size_t lo = count * t / nthreads, hi = count * (t + 1) / nthreads;
      for (size_t i = lo; i < hi; ++i) {
        uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        alg.pre_insert(a, b);
        insert(t, a, b, spare);
        insert(t, b, a, spare);
      }
    };
    std::vector<std::thread> th;
    for (size_t t = 0; t < nthreads; ++t) th.emplace_back(task, (int)t);
    for (auto& x : th) x.join();
  }
  void flush() {  // prep_query: force_flush + flush_workers
    if (alg.shared_dsu_valid) return;
    std::atomic<uint32_t> next{0};
    auto task = [&](int t) {
      while (true) {
        uint32_t v0 = next.fetch_add(256);
        if (v0 >= alg.n) break;
        for (uint32_t v = v0; v < std::min<uint32_t>(alg.n, v0 + 256); ++v) {
          if (gut[v].empty()) continue;
          total += gut[v].size();
          alg.apply_update_batch(t, v, gut[v].data(), gut[v].size());
          gut[v].clear();
        }
      }
    };
    std::vector<std::thread> th;
    for (size_t t = 0; t < nthreads; ++t) th.emplace_back(task, (int)t);
    for (auto& x : th) x.join();
  }
};

}  // namespace oracle

// =============================================================================================
// C ABI used by the Python tests (ctypes).  Handles carry the checksum width.
// =============================================================================================
using namespace oracle;

struct OrAlg {
  int gbytes;
  std::unique_ptr<CCAlg<uint32_t>> a4;
  std::unique_ptr<CCAlg<uint64_t>> a8;
};
struct OrSketch {
  int gbytes;
  std::unique_ptr<Sketch<uint32_t>> s4;
  std::unique_ptr<Sketch<uint64_t>> s8;
};

#define DISPATCH_ALG(h, ...)   \
  do {                         \
    if ((h)->gbytes == 4) {    \
      auto& A = *(h)->a4;      \
      __VA_ARGS__;             \
    } else {                   \
      auto& A = *(h)->a8;      \
      __VA_ARGS__;             \
    }                          \
  } while (0)
#define DISPATCH_SK(h, ...)    \
  do {                         \
    if ((h)->gbytes == 4) {    \
      auto& S = *(h)->s4;      \
      __VA_ARGS__;             \
    } else {                   \
      auto& S = *(h)->s8;      \
      __VA_ARGS__;             \
    }                          \
  } while (0)

extern "C" {

uint64_t or_xxh3_len8(uint64_t input, uint64_t seed) { return xxh3_len8(input, seed); }
void or_xxh3_len8_many(const uint64_t* in, const uint64_t* seeds, size_t cnt, uint64_t* out) {
  for (size_t i = 0; i < cnt; ++i) out[i] = xxh3_len8(in[i], seeds[i]);
}
uint64_t or_calc_vector_length(uint32_t n) { return calc_vector_length(n); }
size_t or_calc_cc_samples(uint32_t n, double f) { return calc_cc_samples(n, f); }
size_t or_calc_bkt_per_col(uint64_t len) { return calc_bkt_per_col(len); }

// ---- sketch ----
void* or_sketch_create(uint64_t vector_len, uint64_t seed, size_t samples, size_t cols, int gbytes) {
  auto* h = new OrSketch{gbytes, nullptr, nullptr};
  if (gbytes == 4) h->s4.reset(new Sketch<uint32_t>(vector_len, seed, samples, cols));
  else h->s8.reset(new Sketch<uint64_t>(vector_len, seed, samples, cols));
  return h;
}
void or_sketch_destroy(void* p) { delete (OrSketch*)p; }
void or_sketch_update_many(void* p, const uint64_t* idx, size_t cnt) {
  auto* h = (OrSketch*)p;
  DISPATCH_SK(h, for (size_t i = 0; i < cnt; ++i) S.update(idx[i]));
}
int or_sketch_sample(void* p, uint64_t* idx, int* result) {
  auto* h = (OrSketch*)p;
  try {
    DISPATCH_SK(h, { auto r = S.sample(); *idx = r.idx; *result = r.result; });
  } catch (OutOfSamples&) {
    return 2;
  }
  return 0;
}
void or_sketch_merge(void* dst, void* src) {
  auto* a = (OrSketch*)dst;
  auto* b = (OrSketch*)src;
  if (a->gbytes == 4) a->s4->merge(*b->s4); else a->s8->merge(*b->s8);
}
void or_sketch_range_merge(void* dst, void* src, size_t start, size_t cnt) {
  auto* a = (OrSketch*)dst;
  auto* b = (OrSketch*)src;
  if (a->gbytes == 4) a->s4->range_merge(*b->s4, start, cnt); else a->s8->range_merge(*b->s8, start, cnt);
}
void or_sketch_merge_raw(void* p, const void* raw) {
  auto* h = (OrSketch*)p;
  if (h->gbytes == 4) h->s4->merge_raw((const Bucket<uint32_t>*)raw);
  else h->s8->merge_raw((const Bucket<uint64_t>*)raw);
}
void or_sketch_zero(void* p) { auto* h = (OrSketch*)p; DISPATCH_SK(h, S.zero_contents()); }
void or_sketch_reset_sample_state(void* p) { auto* h = (OrSketch*)p; DISPATCH_SK(h, S.sample_idx = 0); }
size_t or_sketch_bytes(void* p) { auto* h = (OrSketch*)p; size_t r = 0; DISPATCH_SK(h, r = S.bytes()); return r; }
size_t or_sketch_num_buckets(void* p) { auto* h = (OrSketch*)p; size_t r = 0; DISPATCH_SK(h, r = S.num_buckets); return r; }
size_t or_sketch_bkt_per_col(void* p) { auto* h = (OrSketch*)p; size_t r = 0; DISPATCH_SK(h, r = S.bkt_per_col); return r; }
void or_sketch_serialize(void* p, void* out) {
  auto* h = (OrSketch*)p;
  DISPATCH_SK(h, std::memcpy(out, S.b.data(), S.bytes()));
}

// ---- algorithm ----
void* or_alg_create(uint32_t n, uint64_t seed, double sf, double bf, int gbytes, int eager) {
  auto* h = new OrAlg{gbytes, nullptr, nullptr};
  if (gbytes == 4) h->a4.reset(new CCAlg<uint32_t>(n, seed, sf, bf, eager != 0));
  else h->a8.reset(new CCAlg<uint64_t>(n, seed, sf, bf, eager != 0));
  return h;
}
void or_alg_destroy(void* p) { delete (OrAlg*)p; }
size_t or_alg_sketch_bytes(void* p) { auto* h = (OrAlg*)p; size_t r = 0; DISPATCH_ALG(h, r = A.sk[0]->bytes()); return r; }
size_t or_alg_desired_updates_per_batch(void* p) {
  auto* h = (OrAlg*)p; size_t r = 0; DISPATCH_ALG(h, r = A.desired_updates_per_batch()); return r;
}
void or_alg_allocate_worker_memory(void* p, size_t w) { auto* h = (OrAlg*)p; DISPATCH_ALG(h, A.allocate_worker_memory(w)); }
// Returns 1 when thr has no worker delta (allocate_worker_memory was not called for it).
int or_alg_apply_update_batch(void* p, int thr, uint32_t src, const uint32_t* dsts, size_t k) {
  auto* h = (OrAlg*)p;
  int bad = 0;
  DISPATCH_ALG(h, if (thr < 0 || (size_t)thr >= A.delta.size()) bad = 1; else A.apply_update_batch(thr, src, dsts, k));
  return bad;
}
void or_alg_apply_raw(void* p, uint32_t src, const void* raw) {
  auto* h = (OrAlg*)p;
  if (h->gbytes == 4) h->a4->apply_raw(src, (const Bucket<uint32_t>*)raw);
  else h->a8->apply_raw(src, (const Bucket<uint64_t>*)raw);
}
void or_alg_pre_insert(void* p, uint32_t a, uint32_t b) { auto* h = (OrAlg*)p; DISPATCH_ALG(h, A.pre_insert(a, b)); }
void or_alg_update(void* p, uint32_t a, uint32_t b) { auto* h = (OrAlg*)p; DISPATCH_ALG(h, A.update(a, b)); }
// Apply a whole stream of (src,dst) pairs: both directions, grouped per vertex in stream order
// and applied through apply_update_batch in chunks of at most desired_updates_per_batch().
void or_alg_ingest_pairs(void* p, const uint32_t* pairs, size_t count, int with_pre_insert) {
  auto* h = (OrAlg*)p;
  DISPATCH_ALG(h, {
    if (A.delta.empty()) A.allocate_worker_memory(1);
    size_t gs = std::max<size_t>(1, A.desired_updates_per_batch());
    std::vector<std::vector<uint32_t>> g(A.n);
    for (size_t i = 0; i < count; ++i) {
      uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
      if (with_pre_insert) A.pre_insert(a, b);
      g[a].push_back(b);
      if (g[a].size() == gs) { A.apply_update_batch(0, a, g[a].data(), gs); g[a].clear(); }
      g[b].push_back(a);
      if (g[b].size() == gs) { A.apply_update_batch(0, b, g[b].data(), gs); g[b].clear(); }
    }
    for (uint32_t v = 0; v < A.n; ++v)
      if (!g[v].empty()) A.apply_update_batch(0, v, g[v].data(), g[v].size());
  });
}
int or_alg_has_cached_query(void* p) { auto* h = (OrAlg*)p; int r = 0; DISPATCH_ALG(h, r = A.shared_dsu_valid ? 1 : 0); return r; }
int or_alg_connected_components(void* p, uint32_t* parent_out, uint32_t* ncc) {
  auto* h = (OrAlg*)p;
  try {
    DISPATCH_ALG(h, *ncc = A.connected_components(parent_out));
  } catch (OutOfSamples&) {
    return 2;
  }
  return 0;
}
int or_alg_point_query(void* p, uint32_t a, uint32_t b, int* out) {
  auto* h = (OrAlg*)p;
  try {
    DISPATCH_ALG(h, *out = A.point_query(a, b) ? 1 : 0);
  } catch (OutOfSamples&) {
    return 2;
  }
  return 0;
}
size_t or_alg_last_query_rounds(void* p) { auto* h = (OrAlg*)p; size_t r = 0; DISPATCH_ALG(h, r = A.last_query_rounds); return r; }
// Spanning forest edges (SpanningForest, return_types.cpp:32-41): writes up to cap pairs.
size_t or_alg_spanning_forest(void* p, uint32_t* out, size_t cap) {
  auto* h = (OrAlg*)p;
  size_t k = 0;
  DISPATCH_ALG(h, {
    for (uint32_t s = 0; s < A.n; ++s)
      for (uint32_t d : A.sf[s]) {
        if (k < cap) { out[2 * k] = s; out[2 * k + 1] = d; }
        ++k;
      }
  });
  return k;
}
// Byte image of all vertex sketches in vertex order (what write_binary emits after its header,
// cc_sketch_alg.cpp:614-623).
void or_alg_dump(void* p, uint32_t v0, uint32_t count, void* out) {
  auto* h = (OrAlg*)p;
  DISPATCH_ALG(h, {
    size_t sb = A.sk[0]->bytes();
    for (uint32_t i = 0; i < count; ++i) std::memcpy((char*)out + (size_t)i * sb, A.sk[v0 + i]->b.data(), sb);
  });
}
// CCSketchAlg::write_binary, src/cc_sketch_alg.cpp:614-623: seed (size_t), num_vertices
// (node_id_t), sketches_factor (double), then every vertex's Sketch::serialize image.  Returns 0, or
// 1 when the file cannot be written.
int or_alg_write_binary(void* p, const char* path) {
  auto* h = (OrAlg*)p;
  FILE* f = std::fopen(path, "wb");
  if (!f) return 1;
  bool ok = true;
  DISPATCH_ALG(h, {
    const uint64_t seed = A.seed;
    const uint32_t nv = A.n;
    const double sf = A.sketches_factor;
    ok = std::fwrite(&seed, 8, 1, f) == 1 && std::fwrite(&nv, 4, 1, f) == 1 && std::fwrite(&sf, 8, 1, f) == 1;
    for (uint32_t v = 0; v < nv && ok; ++v) ok = std::fwrite(A.sk[v]->b.data(), A.sk[v]->bytes(), 1, f) == 1;
  });
  return (std::fclose(f) == 0 && ok) ? 0 : 1;
}
// CCSketchAlg::construct_from_serialized_data + the stream constructor, src/cc_sketch_alg.cpp:30-64:
// the header's sketches_factor replaces the configured one, each sketch is read back, and the DSU
// starts invalid (no cached answer).  Returns nullptr when the file is missing or truncated.
void* or_alg_construct_from_serialized(const char* path, double bf, int gbytes, int eager) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return nullptr;
  uint64_t seed = 0;
  uint32_t nv = 0;
  double sf = 1;
  bool ok = std::fread(&seed, 8, 1, f) == 1 && std::fread(&nv, 4, 1, f) == 1 && std::fread(&sf, 8, 1, f) == 1;
  if (!ok || nv < 2) { std::fclose(f); return nullptr; }
  auto* h = (OrAlg*)or_alg_create(nv, seed, sf, bf, gbytes, eager);
  DISPATCH_ALG(h, {
    for (uint32_t v = 0; v < nv && ok; ++v) ok = std::fread(A.sk[v]->b.data(), A.sk[v]->bytes(), 1, f) == 1;
    A.dsu_valid = false;
    A.shared_dsu_valid = false;
  });
  std::fclose(f);
  if (!ok) { or_alg_destroy(h); return nullptr; }
  return h;
}
uint32_t or_exact_cc(uint32_t n, const uint32_t* pairs, size_t count, uint32_t* labels) {
  return exact_cc(n, pairs, count, labels);
}

}  // extern "C"

// CPU baseline over a request trace, the way tools/process_stream.cpp:95-139 times the reference:
// the stream is processed up to each query position (process_stream_until), then prep_query +
// the query is timed as that request's latency.  kinds[i] = 0 for a u-v point query on
// (qa[i], qb[i]), 1 for connected_components().  wall_s = ingest start -> end of the last
// request; lat_s[i] = per-request latency; answers[i] = connected flag or component count.
template <class G>
static int baseline_trace(CCAlg<G>& A, const uint32_t* pairs, size_t count, const uint64_t* qpos,
                          const int* kinds, const uint32_t* qa, const uint32_t* qb, size_t nq,
                          int nthreads, double* wall_s, double* lat_s, uint32_t* answers) {
  CpuDriver<G> drv(A, (size_t)nthreads);
  auto t0 = std::chrono::steady_clock::now();
  size_t done = 0;
  try {
    for (size_t q = 0; q < nq; ++q) {
      size_t upto = std::min<size_t>(qpos[q], count);
      if (upto > done) { drv.process(pairs + 2 * done, upto - done); done = upto; }
      auto q0 = std::chrono::steady_clock::now();
      drv.flush();
      if (kinds[q] == 0) answers[q] = A.point_query(qa[q], qb[q]) ? 1 : 0;
      else answers[q] = A.connected_components(nullptr);
      lat_s[q] = std::chrono::duration<double>(std::chrono::steady_clock::now() - q0).count();
    }
    if (done < count) { drv.process(pairs + 2 * done, count - done); }
    drv.flush();
  } catch (OutOfSamples&) {
    return 2;
  }
  *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

extern "C" int or_baseline_trace(void* p, const uint32_t* pairs, size_t count, const uint64_t* qpos,
                                 const int* kinds, const uint32_t* qa, const uint32_t* qb, size_t nq,
                                 int nthreads, double* wall_s, double* lat_s, uint32_t* answers) {
  auto* h = (OrAlg*)p;
  if (h->gbytes == 4) return baseline_trace(*h->a4, pairs, count, qpos, kinds, qa, qb, nq, nthreads, wall_s, lat_s, answers);
  return baseline_trace(*h->a8, pairs, count, qpos, kinds, qa, qb, nq, nthreads, wall_s, lat_s, answers);
}
