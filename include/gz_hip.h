/* gz_hip.h -- C ABI of the MI355X-native streaming connected-components library (libgz_hip.so).
 *
 * This is the drop-in boundary for GraphZeppelin's hot path.  Each entry point names the reference
 * interface it replaces (paths relative to the reference repository root).  Handles are opaque;
 * every function returns a gz_status and never lets a C++ exception cross the boundary;
 * gz_last_error() gives the calling thread's last message.  Pointers are host pointers unless the
 * name says "device".  No HIP or torch types appear here.
 */
#ifndef GZ_HIP_H
#define GZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status / enums ------------------------------------------------------------------------ */
typedef enum {
  GZ_OK = 0,
  GZ_ERR_INVALID_ARG = 1,
  GZ_ERR_OUT_OF_SAMPLES = 2,  /* OutOfSamplesException, include/sketch.h:193-204 */
  GZ_ERR_UPDATE_LOCKED = 3,   /* UpdateLockedException, include/cc_sketch_alg.h:25-29 */
  GZ_ERR_OOM = 4,             /* device allocation failed */
  GZ_ERR_DEVICE = 5,          /* HIP runtime error, or the HIP library is unusable */
  GZ_ERR_ADMISSION = 6,       /* serving: tenant footprint does not fit the memory budget */
  GZ_ERR_IO = 7               /* file could not be read or written */
} gz_status;

/* SampleResult, include/sketch.h:20-24 (same numbering) */
typedef enum { GZ_GOOD = 0, GZ_ZERO = 1, GZ_FAIL = 2 } gz_sample_result;
/* QueryCode, include/cc_sketch_alg.h:60-63 */
typedef enum { GZ_CONNECTIVITY = 0, GZ_KSPANNINGFORESTS = 1 } gz_query_code;
/* UpdateType (StreamingUtilities, used at include/types.h:12-15): INSERT = 0, DELETE = 1 */
typedef enum { GZ_INSERT = 0, GZ_DELETE = 1, GZ_BREAKPOINT = 2 } gz_update_type;

typedef struct gz_sketch gz_sketch;
typedef struct gz_alg gz_alg;

/* Message of the calling thread's last failed call ("" if none). */
const char* gz_last_error(void);
/* Number of visible HIP devices (0 when no GPU is usable). */
int gz_device_count(int* out);

/* ---- geometry: Sketch::calc_vector_length / calc_cc_samples / calc_bkt_per_col
 *      (include/sketch.h:62-64, 74-76, 181) ---------------------------------------------------- */
uint64_t gz_calc_vector_length(uint32_t num_vertices);
size_t gz_calc_cc_samples(uint32_t num_vertices, double sketches_factor);
size_t gz_calc_bkt_per_col(uint64_t vector_len);

/* ---- Sketch (include/sketch.h:40-191, src/sketch.cpp) --------------------------------------- */
/* Sketch::Sketch(vector_len, seed, num_samples, cols_per_sample), src/sketch.cpp:8-21.
 * gamma_bytes = sizeof(vec_hash_t): 4 (12-byte packed bucket) or 8 (16-byte bucket). */
int gz_sketch_create(uint64_t vector_len, uint64_t seed, size_t num_samples, size_t cols_per_sample,
                     int gamma_bytes, gz_sketch** out);
int gz_sketch_destroy(gz_sketch* s);
/* Sketch::update, src/sketch.cpp:69-83 (one index, or a batch applied in one launch) */
int gz_sketch_update(gz_sketch* s, uint64_t update_idx);
int gz_sketch_update_batch(gz_sketch* s, const uint64_t* idx, size_t count);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): Sketch::update over `count`
 * indices already in device memory, with the kernel and the result of gz_sketch_update_batch, and
 * no allocation or staging.  The kernel's tile descriptors live in a static device array of 1 MiB
 * per GPU, so count is at most 2^26 (GZ_ERR_INVALID_ARG above that).  Every call on a GPU shares
 * that array: calls that overlap on one GPU race.  Returns when the update is done. */
int gz_sketch_update_batch_device_noalloc(gz_sketch* s, const uint64_t* idx_device, size_t count);
/* Sketch::sample, src/sketch.cpp:94-116 -> *result is a gz_sample_result */
int gz_sketch_sample(gz_sketch* s, uint64_t* idx, int* result);
/* Sketch::merge / range_merge / merge_raw_bucket_buffer, src/sketch.cpp:149-186 */
int gz_sketch_merge(gz_sketch* dst, const gz_sketch* src);
int gz_sketch_range_merge(gz_sketch* dst, const gz_sketch* src, size_t start_sample, size_t n_samples);
int gz_sketch_merge_raw_bucket_buffer(gz_sketch* s, const void* raw_buckets);
/* Sketch::zero_contents / reset_sample_state, src/sketch.cpp:86-92, include/sketch.h:167-169 */
int gz_sketch_zero_contents(gz_sketch* s);
int gz_sketch_reset_sample_state(gz_sketch* s);
/* Sketch::serialize / deserialising constructor, src/sketch.cpp:188-190, 23-35: the packed
 * {alpha, gamma} byte image of bucket_array_bytes() bytes */
int gz_sketch_serialize(const gz_sketch* s, void* out, size_t out_bytes);
int gz_sketch_deserialize(gz_sketch* s, const void* in, size_t in_bytes);
/* operator==, src/sketch.cpp:192-204 */
int gz_sketch_equal(const gz_sketch* a, const gz_sketch* b, int* equal);
/* getters, include/sketch.h:164-181 */
size_t gz_sketch_bucket_array_bytes(const gz_sketch* s);
size_t gz_sketch_get_buckets(const gz_sketch* s);
size_t gz_sketch_get_columns(const gz_sketch* s);
size_t gz_sketch_get_num_samples(const gz_sketch* s);
size_t gz_sketch_get_bkt_per_col(const gz_sketch* s);
uint64_t gz_sketch_get_seed(const gz_sketch* s);

/* ---- CCSketchAlg (include/cc_sketch_alg.h:67-257, src/cc_sketch_alg.cpp) -------------------- */
typedef struct {
  double sketches_factor;    /* CCAlgConfiguration::sketches_factor, include/cc_alg_configuration.h */
  double batch_factor;       /* CCAlgConfiguration::batch_factor */
  int gamma_bytes;           /* sizeof(vec_hash_t): 4 or 8 */
  int eager_dsu;             /* 1 = eager DSU (default build), 0 = NO_EAGER_DSU behaviour */
  int batched;               /* 1 = batched form (staged, many vertices per launch);
                                0 = plain form (one launch per apply_update_batch call) */
  uint32_t tile;             /* destinations per update tile; 0 = default */
  uint64_t flush_threshold;  /* device-resident updates buffered before a flush; 0 = default */
} gz_alg_config;
void gz_alg_config_default(gz_alg_config* cfg);

/* CCSketchAlg::CCSketchAlg(num_vertices, seed, config), src/cc_sketch_alg.cpp:11-28 */
int gz_alg_create(uint32_t num_vertices, uint64_t seed, const gz_alg_config* cfg, gz_alg** out);
int gz_alg_destroy(gz_alg* a);
/* allocate_worker_memory / get_desired_updates_per_batch / get_num_vertices,
 * include/cc_sketch_alg.h:146-169, 254 */
int gz_alg_allocate_worker_memory(gz_alg* a, size_t num_workers);
size_t gz_alg_get_desired_updates_per_batch(const gz_alg* a);
uint32_t gz_alg_get_num_vertices(const gz_alg* a);
uint64_t gz_alg_get_seed(const gz_alg* a);
size_t gz_alg_max_rounds(const gz_alg* a);
size_t gz_alg_sketch_bytes(const gz_alg* a); /* bucket_array_bytes() of one vertex sketch */
/* pre_insert(GraphUpdate, thr_id), src/cc_sketch_alg.cpp:79-104 (eager DSU).  Thread-safe. */
int gz_alg_pre_insert(gz_alg* a, uint32_t src, uint32_t dst, int type, int thr_id);
/* apply_update_batch(thr_id, src, dsts), src/cc_sketch_alg.cpp:106-118.  Thread-safe; `dsts` is
 * borrowed for the call only (it is copied before the call returns). */
int gz_alg_apply_update_batch(gz_alg* a, int thr_id, uint32_t src, const uint32_t* dsts, size_t count);
/* apply_raw_buckets_update(src, raw_buckets), src/cc_sketch_alg.cpp:120-123 (packed image) */
int gz_alg_apply_raw_buckets_update(gz_alg* a, uint32_t src, const void* raw_buckets);
/* update(GraphUpdate), src/cc_sketch_alg.cpp:127-133 */
int gz_alg_update(gz_alg* a, uint32_t src, uint32_t dst, int type);
/* Device-resident stream ingest (the batched form of the driver's stream threads + gutters,
 * include/graph_sketch_driver.h:154-178): `count` updates given as (src,dst) uint32 pairs in device
 * memory, each applied to both endpoints' sketches and passed through pre_insert. */
int gz_alg_ingest_pairs_device(gz_alg* a, const uint32_t* pairs_device, size_t count);
/* The same for (src,dst) pairs in host memory.  The copy to the device runs on a stream of its own
 * and overlaps the kernels of earlier batches; pinned memory (hipHostMalloc) is copied directly,
 * pageable memory through pinned staging buffers.  `pairs` is borrowed for the call only: the call
 * returns once the pairs have been read. */
int gz_alg_ingest_pairs(gz_alg* a, const uint32_t* pairs, size_t count);
/* has_cached_query(query_code), include/cc_sketch_alg.h:184-190 -> *out 0/1 */
int gz_alg_has_cached_query(gz_alg* a, int query_code, int* out);
/* prep_query(query_code): the driver's force_flush + flush_workers (include/graph_sketch_driver.h:
 * 193-203): applies every buffered update unless the query is cached.  Blocks until done. */
int gz_alg_prep_query(gz_alg* a, int query_code);
/* connected_components(), src/cc_sketch_alg.cpp:515-557 + ConnectedComponents
 * (src/return_types.cpp:5-19): parent_out (may be NULL) gets n DSU roots, *num_cc the count. */
int gz_alg_connected_components(gz_alg* a, uint32_t* parent_out, uint32_t* num_cc);
/* point_query(a, b), src/cc_sketch_alg.cpp:570-612 -> *connected 0/1 */
int gz_alg_point_query(gz_alg* a, uint32_t u, uint32_t v, int* connected);
/* Many point queries answered against one consistent state (one Borůvka at most). */
int gz_alg_point_queries(gz_alg* a, const uint32_t* us, const uint32_t* vs, size_t count, uint8_t* connected);
/* calc_spanning_forest(), src/cc_sketch_alg.cpp:559-568 + SpanningForest (return_types.cpp:32-41):
 * writes up to `cap` edges as (src,dst) pairs with src < dst; *num_edges = forest size. */
int gz_alg_calc_spanning_forest(gz_alg* a, uint32_t* edges_out, size_t cap, size_t* num_edges);
/* write_binary / construct_from_serialized_data, src/cc_sketch_alg.cpp:614-623, 30-64:
 * seed (8 B) | num_vertices (4 B) | sketches_factor (8 B double) | n packed bucket arrays */
int gz_alg_write_binary(gz_alg* a, const char* path);
int gz_alg_construct_from_serialized_data(const char* path, const gz_alg_config* cfg, gz_alg** out);
/* Packed byte image of `count` vertex sketches starting at v0 (what write_binary emits per vertex,
 * Sketch::serialize): count * gz_alg_sketch_bytes() bytes.  Applies buffered updates first. */
int gz_alg_dump_sketches(gz_alg* a, uint32_t v0, uint32_t count, void* out, size_t out_bytes);
/* Order-sensitive 64-bit digest of every vertex's bucket array (no reference counterpart: a cheap
 * whole-state check).  With pos = v * get_buckets() + b over the packed image,
 *   digest = sum_{v,b} fmix64(x + pos),  x = alpha ^ fmix64(gamma ^ pos * 0x9E3779B97F4A7C15)
 * (mod 2^64, fmix64 = MurmurHash3's 64-bit finaliser).  Applies buffered updates first. */
int gz_alg_digest(gz_alg* a, uint64_t* out);
/* time hooks / counters: last_query_rounds (include/cc_sketch_alg.h:252), updates applied */
size_t gz_alg_last_query_rounds(const gz_alg* a);
uint64_t gz_alg_get_total_updates(const gz_alg* a);
/* Kernel timing (no reference counterpart; the reference only has steady_clock hooks,
 * include/cc_sketch_alg.h:250-252).  When enabled, HIP events on the instance's stream bracket every
 * sketch-update launch and every Borůvka round; alg_bytes is the algorithmic HBM traffic of those
 * launches: per vertex batch of k destinations 4k + 2*min(S, k*(C+1)*B) with S the sketch bytes,
 * C the columns and B the bucket bytes; per Borůvka query n*(min(D,15)+1)*B + 8n for round 0 and
 * e_r*((D+1)*B + 8) for round r >= 1, e_r the vertices round r regroups. */
typedef struct {
  uint64_t launches;
  double total_ms;
  double alg_bytes;
  uint64_t items; /* directed updates (update kernel) or rounds (Borůvka) */
} gz_kernel_stat;
typedef struct {
  gz_kernel_stat update;
  gz_kernel_stat boruvka;
} gz_alg_profile;
int gz_alg_set_profiling(gz_alg* a, int enable);
int gz_alg_get_profile(gz_alg* a, gz_alg_profile* out);
/* No counterpart in the reference (a device execution knob).  enable != 0 (the default): a query
 * replays DSU reset, round 0 and the first speculative Borůvka rounds from a HIP graph built on
 * the instance's first query; 0: the same kernels are launched one by one.  Results are identical. */
int gz_alg_set_query_graphs(gz_alg* a, int enable);
/* No counterpart in the reference (diagnostics).  For each Borůvka round 1..last_query_rounds of
 * the last query that ran Borůvka: the vertices regrouped (those outside a giant supernode) and
 * the giant's root (0xFFFFFFFF when none).  Returns the number of rounds written (<= cap). */
size_t gz_alg_last_query_round_sizes(gz_alg* a, uint32_t* entries, uint32_t* giant_root, size_t cap);
/* No counterpart in the reference (a device execution knob).  enable != 0 (the default): once a
 * giant supernode exists and at most 2048 vertices remain outside it, one single-workgroup launch
 * runs every remaining Borůvka round; 0: every round runs as its own launches.  Results are
 * identical. */
int gz_alg_set_round_finish(gz_alg* a, int enable);
/* No counterpart in the reference (a device execution knob).  enable != 0 (the default): a
 * Borůvka reduce reads a member's records of the round's column that lie past the first line of
 * the column's overflow run only when the member's deep flag for that column is set (rows with
 * 4-byte checksums; every write of such a record nonzero sets the flag, and most vertices never
 * touch those depths); 0: it reads every record.  Results are identical. */
int gz_alg_set_deep_skip(gz_alg* a, int enable);
/* No counterpart in the reference (diagnostics).  The Borůvka round from which the last query ran
 * its remaining rounds in the single-workgroup finish, or 0 when it did not. */
uint32_t gz_alg_last_query_finish_round(const gz_alg* a);
/* No counterpart in the reference (the update-tile and flush-threshold autotuner).  Times the
 * device ingest of a sample stream (count undirected updates as (src,dst) pairs in device memory,
 * as gz_alg_ingest_pairs_device takes them) on scratch instances with num_vertices vertices and the
 * base configuration, for every candidate tile size (destinations per update wave, so also waves
 * per vertex) and flush threshold (updates staged on the device per grouping and update launch).
 * *out gets the base configuration with the fastest pair; the report lists every trial.  The
 * caller's instances are not touched, and results never depend on the choice: it changes how the
 * work is cut into launches, not what is computed.  Needs device memory for one more instance. */
#define GZ_AUTOTUNE_MAX_TRIALS 16
typedef struct {
  uint32_t trials;
  uint32_t tile[GZ_AUTOTUNE_MAX_TRIALS];
  uint64_t flush_threshold[GZ_AUTOTUNE_MAX_TRIALS];
  double ms[GZ_AUTOTUNE_MAX_TRIALS];  /* best of two timed ingests of the sample */
} gz_autotune_report;
int gz_autotune(uint32_t num_vertices, uint64_t seed, const gz_alg_config* base, const uint32_t* pairs_device,
                size_t count, gz_alg_config* out, gz_autotune_report* report);
/* Device footprint in bytes of an instance with these parameters (sketch slab + work buffers). */
uint64_t gz_alg_footprint_bytes(uint32_t num_vertices, const gz_alg_config* cfg);

#ifdef __cplusplus
}
#endif
#endif /* GZ_HIP_H */
