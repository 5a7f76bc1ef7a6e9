// graphzeppelin_amd.hpp -- header-only C++ facade over libgz_hip.so (include/gz_hip.h) with the
// reference's class names and method signatures, so code written against GraphZeppelin's
// CCSketchAlg / Sketch / GraphSketchDriver compiles against the MI355X library unchanged.
//
//   reference                                   here
//   include/sketch.h          Sketch            gzamd::Sketch           (buckets in HBM)
//   include/cc_sketch_alg.h   CCSketchAlg       gzamd::CCSketchAlg      (all vertex sketches in HBM)
//   include/cc_alg_configuration.h              gzamd::CCAlgConfiguration
//   include/return_types.h    ConnectedComponents, SpanningForest
//   include/graph_sketch_driver.h GraphSketchDriver<Alg>  (same plugin interface)
//   include/driver_configuration.h DriverConfiguration
//   StreamingUtilities        GraphStream, BinaryFileStream, AsciiFileStream, GraphStreamUpdate
//
// The names are also exported at global scope unless GZAMD_NO_GLOBAL_NAMES is defined.
// Link with -lgz_hip.  C++14.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <vector>

#include <fcntl.h>
#include <unistd.h>

#include "gz_hip.h"

namespace gzamd {

typedef uint32_t node_id_t;
typedef uint64_t edge_id_t;
typedef uint64_t vec_t;

struct Edge {
  node_id_t src = 0;
  node_id_t dst = 0;
};
enum UpdateType { INSERT = 0, DELETE = 1, BREAKPOINT = 2 };
struct GraphUpdate {  // include/types.h:12-15
  Edge edge;
  UpdateType type;
};
struct GraphStreamUpdate {  // StreamingUtilities: {Edge edge; uint8 type;}
  Edge edge;
  uint8_t type;
};
enum SampleResult { GOOD = GZ_GOOD, ZERO = GZ_ZERO, FAIL = GZ_FAIL };
struct SketchSample {
  vec_t idx;
  SampleResult result;
};
enum QueryCode { CONNECTIVITY = GZ_CONNECTIVITY, KSPANNINGFORESTS = GZ_KSPANNINGFORESTS };
static constexpr edge_id_t END_OF_STREAM = ~edge_id_t(0);

// ---- exceptions (include/sketch.h:193-204, include/cc_sketch_alg.h:25-29) ----
class GZException : public std::runtime_error {
 public:
  int code;
  GZException(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
class OutOfSamplesException : public GZException {
 public:
  explicit OutOfSamplesException(const std::string& m) : GZException(GZ_ERR_OUT_OF_SAMPLES, m) {}
};
class UpdateLockedException : public GZException {
 public:
  explicit UpdateLockedException(const std::string& m) : GZException(GZ_ERR_UPDATE_LOCKED, m) {}
};
class DriverException : public std::exception {
  std::string err_msg;

 public:
  explicit DriverException(std::string msg) : err_msg(std::move(msg)) {}
  const char* what() const noexcept override { return err_msg.c_str(); }
};

inline void check(int rc) {
  if (rc == GZ_OK) return;
  const std::string m = gz_last_error();
  if (rc == GZ_ERR_OUT_OF_SAMPLES) throw OutOfSamplesException(m);
  if (rc == GZ_ERR_UPDATE_LOCKED) throw UpdateLockedException(m);
  throw GZException(rc, m);
}

inline edge_id_t concat_pairing_fn(node_id_t i, node_id_t j) {  // src/util.cpp:45-51
  if (i > j) std::swap(i, j);
  return ((edge_id_t)i << 32) | j;
}
inline Edge inv_concat_pairing_fn(edge_id_t idx) {  // src/util.cpp:53-57
  return Edge{(node_id_t)(idx >> 32), (node_id_t)(idx & 0xFFFFFFFF)};
}

// ---- Sketch (include/sketch.h:40-191) ----
class Sketch {
  gz_sketch* h_ = nullptr;

 public:
  static vec_t calc_vector_length(node_id_t n) { return gz_calc_vector_length(n); }
  static size_t calc_cc_samples(node_id_t n, double f) { return gz_calc_cc_samples(n, f); }
  static size_t calc_bkt_per_col(size_t len) { return gz_calc_bkt_per_col(len); }

  Sketch(vec_t vector_len, uint64_t seed, size_t num_samples = 1, size_t cols_per_sample = 1, int gamma_bytes = 4) {
    check(gz_sketch_create(vector_len, seed, num_samples, cols_per_sample, gamma_bytes, &h_));
  }
  Sketch(vec_t vector_len, uint64_t seed, std::istream& binary_in, size_t num_samples = 1,
         size_t cols_per_sample = 1, int gamma_bytes = 4)
      : Sketch(vector_len, seed, num_samples, cols_per_sample, gamma_bytes) {
    std::vector<char> buf(bucket_array_bytes());
    binary_in.read(buf.data(), (std::streamsize)buf.size());
    check(gz_sketch_deserialize(h_, buf.data(), buf.size()));
  }
  Sketch(const Sketch&) = delete;
  Sketch& operator=(const Sketch&) = delete;
  ~Sketch() { gz_sketch_destroy(h_); }

  void update(const vec_t idx) { check(gz_sketch_update(h_, idx)); }
  void update_batch(const std::vector<vec_t>& idx) { check(gz_sketch_update_batch(h_, idx.data(), idx.size())); }
  SketchSample sample() {
    uint64_t i = 0;
    int r = 0;
    check(gz_sketch_sample(h_, &i, &r));
    return SketchSample{i, (SampleResult)r};
  }
  void merge(const Sketch& o) { check(gz_sketch_merge(h_, o.h_)); }
  void range_merge(const Sketch& o, size_t start_sample, size_t n_samples) {
    check(gz_sketch_range_merge(h_, o.h_, start_sample, n_samples));
  }
  void merge_raw_bucket_buffer(const void* raw) { check(gz_sketch_merge_raw_bucket_buffer(h_, raw)); }
  void zero_contents() { check(gz_sketch_zero_contents(h_)); }
  void reset_sample_state() { check(gz_sketch_reset_sample_state(h_)); }
  void serialize(std::ostream& out) const {
    std::vector<char> buf(bucket_array_bytes());
    check(gz_sketch_serialize(h_, buf.data(), buf.size()));
    out.write(buf.data(), (std::streamsize)buf.size());
  }
  friend bool operator==(const Sketch& a, const Sketch& b) {
    int e = 0;
    check(gz_sketch_equal(a.h_, b.h_, &e));
    return e != 0;
  }
  size_t bucket_array_bytes() const { return gz_sketch_bucket_array_bytes(h_); }
  uint64_t get_seed() const { return gz_sketch_get_seed(h_); }
  size_t column_seed(size_t c) const { return get_seed() + c * 5; }
  size_t checksum_seed() const { return get_seed(); }
  size_t get_columns() const { return gz_sketch_get_columns(h_); }
  size_t get_buckets() const { return gz_sketch_get_buckets(h_); }
  size_t get_num_samples() const { return gz_sketch_get_num_samples(h_); }
};

// ---- configuration (include/cc_alg_configuration.h) ----
class CCAlgConfiguration {
 public:
  gz_alg_config c;
  CCAlgConfiguration() { gz_alg_config_default(&c); }
  CCAlgConfiguration& disk_dir(const std::string&) { return *this; }  // no disk tier on the device path
  CCAlgConfiguration& sketches_factor(double f) {
    c.sketches_factor = f > 0 ? f : 1;
    return *this;
  }
  CCAlgConfiguration& batch_factor(double f) {
    c.batch_factor = f > 0 ? f : 1;
    return *this;
  }
  CCAlgConfiguration& gamma_bytes(int b) {
    c.gamma_bytes = b;
    return *this;
  }
  CCAlgConfiguration& eager_dsu(bool e) {
    c.eager_dsu = e ? 1 : 0;
    return *this;
  }
  CCAlgConfiguration& batched(bool b) {
    c.batched = b ? 1 : 0;
    return *this;
  }
  double get_sketches_factor() const { return c.sketches_factor; }
  double get_batch_factor() const { return c.batch_factor; }
  friend std::ostream& operator<<(std::ostream& out, const CCAlgConfiguration& conf) {
    out << "Connected Components Algorithm Configuration:\n"
        << " Sketching algorithm   = CameoSketch (MI355X)\n"
        << " Using Eager DSU       = " << (conf.c.eager_dsu ? "True" : "False") << "\n"
        << " Num sketches factor   = " << conf.c.sketches_factor << "\n"
        << " Batch size factor     = " << conf.c.batch_factor << "\n"
        << " Checksum bytes        = " << conf.c.gamma_bytes;
    return out;
  }
};

// ---- query results (include/return_types.h) ----
class ConnectedComponents {
  std::vector<node_id_t> parent_arr;
  node_id_t num_cc = 0;

 public:
  ConnectedComponents(std::vector<node_id_t>&& parents, node_id_t ncc) : parent_arr(std::move(parents)), num_cc(ncc) {}
  std::vector<std::set<node_id_t>> get_component_sets() const {
    std::map<node_id_t, std::set<node_id_t>> temp;
    for (node_id_t i = 0; i < parent_arr.size(); ++i) temp[parent_arr[i]].insert(i);
    std::vector<std::set<node_id_t>> out;
    for (auto& kv : temp) out.push_back(std::move(kv.second));
    return out;
  }
  bool is_connected(node_id_t a, node_id_t b) const { return parent_arr[a] == parent_arr[b]; }
  node_id_t size() const { return num_cc; }
};
class SpanningForest {
  std::vector<Edge> edges;

 public:
  explicit SpanningForest(std::vector<Edge>&& e) : edges(std::move(e)) {}
  const std::vector<Edge>& get_edges() const { return edges; }
};

// ---- CCSketchAlg (include/cc_sketch_alg.h) ----
class CCSketchAlg {
  gz_alg* h_ = nullptr;
  node_id_t n_ = 0;
  CCAlgConfiguration config;
  explicit CCSketchAlg(gz_alg* h, const CCAlgConfiguration& c) : h_(h), n_(gz_alg_get_num_vertices(h)), config(c) {}

 public:
  CCSketchAlg(node_id_t num_vertices, size_t seed, CCAlgConfiguration cfg = CCAlgConfiguration())
      : n_(num_vertices), config(cfg) {
    check(gz_alg_create(num_vertices, seed, &config.c, &h_));
  }
  CCSketchAlg(const CCSketchAlg&) = delete;
  CCSketchAlg& operator=(const CCSketchAlg&) = delete;
  ~CCSketchAlg() { gz_alg_destroy(h_); }

  static CCSketchAlg* construct_from_serialized_data(const std::string& input_file,
                                                     CCAlgConfiguration cfg = CCAlgConfiguration()) {
    gz_alg* h = nullptr;
    check(gz_alg_construct_from_serialized_data(input_file.c_str(), &cfg.c, &h));
    return new CCSketchAlg(h, cfg);
  }

  // plugin interface used by GraphSketchDriver (include/graph_sketch_driver.h:23-69)
  void allocate_worker_memory(size_t num_workers) { check(gz_alg_allocate_worker_memory(h_, num_workers)); }
  size_t get_desired_updates_per_batch() { return gz_alg_get_desired_updates_per_batch(h_); }
  node_id_t get_num_vertices() { return n_; }
  void pre_insert(GraphUpdate upd, int thr_id = 0) {
    check(gz_alg_pre_insert(h_, upd.edge.src, upd.edge.dst, (int)upd.type, thr_id));
  }
  void apply_update_batch(int thr_id, node_id_t src, const std::vector<node_id_t>& dsts) {
    check(gz_alg_apply_update_batch(h_, thr_id, src, dsts.data(), dsts.size()));
  }
  bool has_cached_query(int query_code) {
    int r = 0;
    check(gz_alg_has_cached_query(h_, query_code, &r));
    return r != 0;
  }
  void print_configuration() { std::cout << config << std::endl; }
  // Device ingest of whole update chunks (pre_insert + both directions): lets the driver skip the
  // host gutters.  Not part of the reference interface; GraphSketchDriver uses it when present.
  // Device ingest of (src, dst) pairs in host memory (pre_insert + both directions), DMA'd
  // asynchronously; the pairs are borrowed for the call only.
  void ingest_pairs(const uint32_t* pairs, size_t count) { check(gz_alg_ingest_pairs(h_, pairs, count)); }
  void ingest_updates(const GraphStreamUpdate* upd, size_t count) {
    std::vector<uint32_t> pairs(2 * count);
    for (size_t i = 0; i < count; ++i) {
      pairs[2 * i] = upd[i].edge.src;
      pairs[2 * i + 1] = upd[i].edge.dst;
    }
    check(gz_alg_ingest_pairs(h_, pairs.data(), count));
  }
  void prep_query(int query_code) { check(gz_alg_prep_query(h_, query_code)); }

  void apply_raw_buckets_update(node_id_t src, const void* raw) { check(gz_alg_apply_raw_buckets_update(h_, src, raw)); }
  void update(GraphUpdate upd) { check(gz_alg_update(h_, upd.edge.src, upd.edge.dst, (int)upd.type)); }

  ConnectedComponents connected_components() {
    cc_alg_start = std::chrono::steady_clock::now();
    std::vector<node_id_t> par(n_);
    uint32_t ncc = 0;
    check(gz_alg_connected_components(h_, par.data(), &ncc));
    cc_alg_end = std::chrono::steady_clock::now();
    last_query_rounds = gz_alg_last_query_rounds(h_);
    return ConnectedComponents(std::move(par), ncc);
  }
  bool point_query(node_id_t a, node_id_t b) {
    cc_alg_start = std::chrono::steady_clock::now();
    int r = 0;
    check(gz_alg_point_query(h_, a, b, &r));
    cc_alg_end = std::chrono::steady_clock::now();
    last_query_rounds = gz_alg_last_query_rounds(h_);
    return r != 0;
  }
  SpanningForest calc_spanning_forest() {
    std::vector<uint32_t> raw(2 * (size_t)n_);
    size_t k = 0;
    check(gz_alg_calc_spanning_forest(h_, raw.data(), n_, &k));
    std::vector<Edge> e(k);
    for (size_t i = 0; i < k; ++i) e[i] = Edge{raw[2 * i], raw[2 * i + 1]};
    return SpanningForest(std::move(e));
  }
  void write_binary(const std::string& filename) { check(gz_alg_write_binary(h_, filename.c_str())); }
  std::vector<uint8_t> dump_sketches(node_id_t v0, node_id_t count) {
    std::vector<uint8_t> out((size_t)count * gz_alg_sketch_bytes(h_));
    check(gz_alg_dump_sketches(h_, v0, count, out.data(), out.size()));
    return out;
  }

  std::chrono::steady_clock::time_point cc_alg_start, cc_alg_end;
  size_t last_query_rounds = 0;
  size_t get_seed() { return gz_alg_get_seed(h_); }
  size_t max_rounds() { return gz_alg_max_rounds(h_); }
  gz_alg* handle() { return h_; }
};

// ---- graph streams (StreamingUtilities' GraphStream API) ----
class GraphStream {
 public:
  virtual ~GraphStream() = default;
  // Fill up to `n` updates; a BREAKPOINT record marks the break point / end of stream.
  virtual size_t get_update_buffer(GraphStreamUpdate* buf, size_t n) = 0;
  virtual bool set_break_point(edge_id_t break_idx) = 0;
  virtual bool get_update_is_thread_safe() { return true; }
  node_id_t vertices() const { return num_vertices; }
  edge_id_t edges() const { return num_edges; }

 protected:
  node_id_t num_vertices = 0;
  edge_id_t num_edges = 0;
};

// In-memory stream of (src, dst, type) records; thread safe.
class MemoryStream : public GraphStream {
 protected:
  std::vector<GraphStreamUpdate> upd;
  std::mutex mtx;
  edge_id_t pos = 0, brk = END_OF_STREAM;

 public:
  MemoryStream() = default;
  MemoryStream(node_id_t n, std::vector<GraphStreamUpdate> u) : upd(std::move(u)) {
    num_vertices = n;
    num_edges = upd.size();
  }
  size_t get_update_buffer(GraphStreamUpdate* buf, size_t n) override {
    std::lock_guard<std::mutex> lk(mtx);
    const edge_id_t end = std::min<edge_id_t>(brk, num_edges);
    size_t k = 0;
    while (k < n && pos < end) buf[k++] = upd[pos++];
    if (k < n && pos >= end) buf[k++] = GraphStreamUpdate{Edge{}, (uint8_t)BREAKPOINT};
    return k;
  }
  bool set_break_point(edge_id_t idx) override {
    std::lock_guard<std::mutex> lk(mtx);
    if (idx != END_OF_STREAM && idx < pos) return false;
    brk = idx;
    return true;
  }
};

// Binary stream file (README.md:57-72): u32 vertices | u64 updates | 9-byte {u8 type, u32 src, u32 dst}.
// Read from the file as the stream threads ask for updates: each get_update_buffer call reserves
// the next run of records under a lock and reads them with pread, so several reader threads read
// disjoint parts of the file at once (thread safe, like StreamingUtilities' BinaryFileStream).
class BinaryFileStream : public GraphStream {
  int fd = -1;
  std::mutex mtx;
  edge_id_t pos = 0, brk = END_OF_STREAM;
  static constexpr uint64_t kHeader = 12, kRecord = 9;

 public:
  explicit BinaryFileStream(const std::string& path) {
    fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw DriverException("cannot open " + path);
    unsigned char h[kHeader];
    if (::pread(fd, h, kHeader, 0) != (ssize_t)kHeader) {
      ::close(fd);
      throw DriverException("truncated stream header " + path);
    }
    uint32_t n = 0;
    uint64_t m = 0;
    std::memcpy(&n, h, 4);
    std::memcpy(&m, h + 4, 8);
    num_vertices = n;
    num_edges = m;
  }
  BinaryFileStream(const BinaryFileStream&) = delete;
  BinaryFileStream& operator=(const BinaryFileStream&) = delete;
  ~BinaryFileStream() override {
    if (fd >= 0) ::close(fd);
  }
  size_t get_update_buffer(GraphStreamUpdate* buf, size_t n) override {
    edge_id_t start = 0, cnt = 0;
    bool at_end = false;
    {
      std::lock_guard<std::mutex> lk(mtx);
      const edge_id_t end = std::min<edge_id_t>(brk, num_edges);
      start = pos;
      cnt = std::min<edge_id_t>(n, end > pos ? end - pos : 0);
      pos += cnt;
      at_end = pos >= end;
    }
    if (cnt) {
      std::vector<unsigned char> raw(cnt * kRecord);
      size_t got = 0;
      while (got < raw.size()) {
        const ssize_t r = ::pread(fd, raw.data() + got, raw.size() - got, (off_t)(kHeader + start * kRecord + got));
        if (r <= 0) throw DriverException("stream file read failed");
        got += (size_t)r;
      }
      for (edge_id_t i = 0; i < cnt; ++i) {
        const unsigned char* r = &raw[kRecord * i];
        uint32_t a, b;
        std::memcpy(&a, r + 1, 4);
        std::memcpy(&b, r + 5, 4);
        buf[i] = GraphStreamUpdate{Edge{a, b}, r[0]};
      }
    }
    size_t k = (size_t)cnt;
    if (k < n && at_end) buf[k++] = GraphStreamUpdate{Edge{}, (uint8_t)BREAKPOINT};
    return k;
  }
  bool set_break_point(edge_id_t idx) override {
    std::lock_guard<std::mutex> lk(mtx);
    if (idx != END_OF_STREAM && idx < pos) return false;
    brk = idx;
    return true;
  }
};

// ASCII stream: "n m" then "src dst" lines (insert-only) or "type src dst" lines (with_types).
class AsciiFileStream : public MemoryStream {
 public:
  explicit AsciiFileStream(const std::string& path, bool with_types = true) {
    std::ifstream f(path);
    if (!f) throw DriverException("cannot open " + path);
    uint64_t n = 0, m = 0;
    f >> n >> m;
    upd.resize(m);
    for (uint64_t i = 0; i < m; ++i) {
      int t = 0;
      uint32_t s, d;
      if (with_types) f >> t;
      f >> s >> d;
      upd[i] = GraphStreamUpdate{Edge{s, d}, (uint8_t)t};
    }
    num_vertices = (node_id_t)n;
    num_edges = m;
  }
};

// ---- driver (include/graph_sketch_driver.h, include/driver_configuration.h) ----
enum GutterSystem { GUTTERTREE, STANDALONE, CACHETREE, DEVICE };
class DriverConfiguration {
  GutterSystem sys = DEVICE;
  size_t workers = 1;

 public:
  DriverConfiguration& gutter_sys(GutterSystem g) {
    sys = g;
    return *this;
  }
  DriverConfiguration& worker_threads(size_t w) {
    workers = w ? w : 1;
    return *this;
  }
  DriverConfiguration& disk_dir(const std::string&) { return *this; }
  GutterSystem get_gutter_sys() const { return sys; }
  size_t get_worker_threads() const { return workers; }
};

namespace detail {
template <class A, class = void>
struct has_ingest : std::false_type {};
template <class A>
struct has_ingest<A, decltype(std::declval<A&>().ingest_updates((const GraphStreamUpdate*)nullptr, size_t(0)), void())>
    : std::true_type {};
template <class A, class = void>
struct has_ingest_pairs : std::false_type {};
template <class A>
struct has_ingest_pairs<A, decltype(std::declval<A&>().ingest_pairs((const uint32_t*)nullptr, size_t(0)), void())>
    : std::true_type {};
template <class A, class = void>
struct has_prep_query : std::false_type {};
template <class A>
struct has_prep_query<A, decltype(std::declval<A&>().prep_query(0), void())> : std::true_type {};
}  // namespace detail

// GraphSketchDriver<Alg> with the reference's interface.  With DEVICE gutters (the default) and an
// algorithm that offers ingest_updates(), stream chunks go straight to the device, which groups
// them per vertex; otherwise updates take the reference route (pre_insert, two directed inserts into
// per-vertex host gutters of get_desired_updates_per_batch() destinations, batch_callback on full
// gutters, all gutters drained by prep_query).
template <class Alg>
class GraphSketchDriver {
  Alg* alg;
  GraphStream* stream;
  DriverConfiguration conf;
  size_t num_stream_threads;
  static constexpr size_t update_array_size = 4000;  // graph_sketch_driver.h:84
  std::atomic<size_t> total_updates{0};
  std::vector<std::vector<node_id_t>> gutters;
  std::vector<std::mutex> gmtx;
  size_t gutter_size;
  bool device_path;

  void host_insert(node_id_t src, node_id_t dst, int thr) {
    std::vector<node_id_t> full;
    {
      std::lock_guard<std::mutex> lk(gmtx[src]);
      auto& g = gutters[src];
      g.push_back(dst);
      if (g.size() < gutter_size) return;
      full.swap(g);
    }
    batch_callback(thr, src, full);
  }

  // A reader thread's host gutter: its updates as (src, dst) pairs, handed to the device whole.
  static constexpr size_t reader_batch = size_t(1) << 20;
  template <class A = Alg>
  typename std::enable_if<detail::has_ingest_pairs<A>::value>::type device_ingest(const std::vector<uint32_t>& p) {
    alg->ingest_pairs(p.data(), p.size() / 2);
  }
  template <class A = Alg>
  typename std::enable_if<!detail::has_ingest_pairs<A>::value && detail::has_ingest<A>::value>::type device_ingest(
      const std::vector<uint32_t>& p) {
    std::vector<GraphStreamUpdate> u(p.size() / 2);
    for (size_t i = 0; i < u.size(); ++i) u[i] = GraphStreamUpdate{Edge{p[2 * i], p[2 * i + 1]}, (uint8_t)INSERT};
    alg->ingest_updates(u.data(), u.size());
  }
  template <class A = Alg>
  typename std::enable_if<!detail::has_ingest_pairs<A>::value && !detail::has_ingest<A>::value>::type device_ingest(
      const std::vector<uint32_t>&) {}
  template <class A = Alg>
  typename std::enable_if<detail::has_prep_query<A>::value>::type device_flush(int code) {
    alg->prep_query(code);
  }
  template <class A = Alg>
  typename std::enable_if<!detail::has_prep_query<A>::value>::type device_flush(int) {}

 public:
  GraphSketchDriver(Alg* a, GraphStream* s, DriverConfiguration c = DriverConfiguration(), size_t n_stream_threads = 1)
      : alg(a), stream(s), conf(c), num_stream_threads(n_stream_threads), gutters(a->get_num_vertices()),
        gmtx(a->get_num_vertices()) {
    alg->allocate_worker_memory(conf.get_worker_threads());
    gutter_size = std::max<size_t>(1, alg->get_desired_updates_per_batch());
    device_path = conf.get_gutter_sys() == DEVICE && (detail::has_ingest<Alg>::value || detail::has_ingest_pairs<Alg>::value);
    if (num_stream_threads > 1 && !stream->get_update_is_thread_safe()) num_stream_threads = 1;
    alg->print_configuration();
  }

  // graph_sketch_driver.h:141-191
  void process_stream_until(edge_id_t break_edge_idx) {
    if (!stream->set_break_point(break_edge_idx))
      throw DriverException("Could not correctly set breakpoint: " + std::to_string(break_edge_idx));
    std::mutex dev_mtx;
    auto task = [&](int thr) {
      std::vector<GraphStreamUpdate> buf(update_array_size);
      std::vector<uint32_t> gutter;  // device path: this reader's pairs, handed over when full
      if (device_path) gutter.reserve(2 * reader_batch);
      auto hand_over = [&] {
        if (gutter.empty()) return;
        std::lock_guard<std::mutex> lk(dev_mtx);
        device_ingest(gutter);
        total_updates += gutter.size();  // two directed updates per pair
        gutter.clear();
      };
      while (true) {
        size_t k = stream->get_update_buffer(buf.data(), update_array_size);
        size_t m = 0;
        bool stop = false;
        while (m < k && buf[m].type != BREAKPOINT) ++m;
        if (m < k) stop = true;
        if (device_path) {
          for (size_t i = 0; i < m; ++i) {
            gutter.push_back(buf[i].edge.src);
            gutter.push_back(buf[i].edge.dst);
          }
          if (gutter.size() >= 2 * reader_batch) hand_over();
        } else {
          for (size_t i = 0; i < m; ++i) {
            GraphUpdate upd{buf[i].edge, (UpdateType)buf[i].type};
            alg->pre_insert(upd, thr);
            host_insert(upd.edge.src, upd.edge.dst, thr);
            host_insert(upd.edge.dst, upd.edge.src, thr);
          }
        }
        if (stop || k == 0) {
          hand_over();
          return;
        }
      }
    };
    std::vector<std::thread> th;
    for (size_t i = 0; i < num_stream_threads; ++i) th.emplace_back(task, (int)i);
    for (auto& t : th) t.join();
  }

  // graph_sketch_driver.h:193-203
  void prep_query(int query_code) {
    flush_start = flush_end = std::chrono::steady_clock::now();
    if (alg->has_cached_query(query_code)) return;
    if (device_path) device_flush(query_code);  // the device gutters: group and apply every pending update
    if (!device_path)
      for (node_id_t v = 0; v < gutters.size(); ++v)
        if (!gutters[v].empty()) {
          std::vector<node_id_t> full;
          full.swap(gutters[v]);
          batch_callback(0, v, full);
        }
    flush_end = std::chrono::steady_clock::now();
  }

  // graph_sketch_driver.h:205-209
  void batch_callback(int thr_id, node_id_t src, const std::vector<node_id_t>& dsts) {
    total_updates += dsts.size();
    alg->apply_update_batch(thr_id, src, dsts);
  }
  size_t get_total_updates() { return total_updates.load(); }

  std::chrono::steady_clock::time_point flush_start, flush_end;
};

}  // namespace gzamd

#ifndef GZAMD_NO_GLOBAL_NAMES
using gzamd::node_id_t;
using gzamd::edge_id_t;
using gzamd::vec_t;
using gzamd::AsciiFileStream;
using gzamd::BinaryFileStream;
using gzamd::CCAlgConfiguration;
using gzamd::CCSketchAlg;
using gzamd::ConnectedComponents;
using gzamd::DriverConfiguration;
using gzamd::DriverException;
using gzamd::Edge;
using gzamd::GraphSketchDriver;
using gzamd::GraphStream;
using gzamd::GraphStreamUpdate;
using gzamd::GraphUpdate;
using gzamd::MemoryStream;
using gzamd::OutOfSamplesException;
using gzamd::Sketch;
using gzamd::SketchSample;
using gzamd::SpanningForest;
using gzamd::UpdateLockedException;
using gzamd::BREAKPOINT;
using gzamd::CACHETREE;
using gzamd::CONNECTIVITY;
using gzamd::DELETE;
using gzamd::DEVICE;
using gzamd::END_OF_STREAM;
using gzamd::FAIL;
using gzamd::GOOD;
using gzamd::GUTTERTREE;
using gzamd::INSERT;
using gzamd::KSPANNINGFORESTS;
using gzamd::STANDALONE;
using gzamd::ZERO;
using gzamd::concat_pairing_fn;
using gzamd::inv_concat_pairing_fn;
#endif
