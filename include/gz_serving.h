/* gz_serving.h -- request loop over device-resident graph-stream sketches (libgz_hip.so).
 *
 * The reference has no serving layer; its closest counterpart is the measurement loop of
 * tools/process_stream.cpp:95-139 (ingest up to a breakpoint, prep_query + connected_components,
 * timed per query).  These entry points run that loop natively over a request trace, for one
 * stream or for many concurrent streams ("tenants") admitted by sketch footprint.
 */
#ifndef GZ_SERVING_H
#define GZ_SERVING_H

#include <stddef.h>
#include <stdint.h>

#include "gz_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GZ_REQ_UPDATES = 0, /* apply pairs [offset, offset + count) of the tenant's stream */
  GZ_REQ_POINT = 1,   /* point_query(u, v), src/cc_sketch_alg.cpp:570-612 */
  GZ_REQ_CC = 2       /* connected_components(), src/cc_sketch_alg.cpp:515-557 */
} gz_request_kind;

typedef struct {
  uint32_t kind;
  uint32_t tenant;
  uint64_t offset;
  uint64_t count;
  uint32_t u, v;
  double arrival_s; /* multi-tenant replay: seconds after the start at which the request arrives */
} gz_request;

typedef struct {
  double wall_s;            /* first request -> every update applied (process_stream.cpp:130-136) */
  uint64_t updates;         /* stream updates applied */
  uint64_t queries;         /* query requests answered */
  uint64_t boruvka_runs;    /* queries that needed a Borůvka pass */
  uint64_t cached_answers;  /* query groups answered from the eager DSU */
  uint64_t query_groups;    /* groups of queries answered against one state */
} gz_serve_stats;

/* Batched/served form over one algorithm instance.  `pairs` holds the stream as (src,dst) uint32
 * pairs (device memory when pairs_on_device != 0, else host memory); it must stay unchanged until
 * the call returns, since batches are read in place (device memory, pinned host memory) or copied
 * asynchronously (pageable host memory, through pinned staging).  Consecutive queries form one
 * group answered against one state (at most one Borůvka); each request's latency is measured from
 * the moment the loop reaches it to its answer.  latency_s / answers: nreq entries (updates get 0;
 * point queries 0/1; CC queries the component count). */
int gz_alg_serve_trace(gz_alg* a, const uint32_t* pairs, int pairs_on_device, const gz_request* reqs, size_t nreq,
                       double* latency_s, uint32_t* answers, gz_serve_stats* stats);

/* Plain form over an instance created with batched = 0: the reference's ingest shape on one host
 * thread -- pre_insert per update, two directed inserts into per-vertex host gutters of
 * get_desired_updates_per_batch() destinations (include/graph_sketch_driver.h:154-178), one
 * apply_update_batch (one kernel launch) per full gutter, and synchronous queries that first flush
 * every non-empty gutter.  `pairs` in host memory. */
int gz_alg_plain_trace(gz_alg* a, const uint32_t* pairs, const gz_request* reqs, size_t nreq, double* latency_s,
                       uint32_t* answers, gz_serve_stats* stats);

/* ---- multi-tenant server ------------------------------------------------------------------ */
typedef struct gz_server gz_server;

/* hbm_budget_bytes = 0: the device's free memory at creation.  workers: host threads that drive
 * tenants concurrently (each tenant's device work runs on its own HIP stream). */
int gz_server_create(uint64_t hbm_budget_bytes, int workers, gz_server** out);
int gz_server_destroy(gz_server* s);
/* Admit a stream of num_vertices vertices.  Returns GZ_ERR_ADMISSION (and *tenant = -1) when its
 * footprint (gz_alg_footprint_bytes) does not fit in the unreserved budget. */
int gz_server_admit(gz_server* s, uint32_t num_vertices, uint64_t seed, const gz_alg_config* cfg, int* tenant);
int gz_server_release(gz_server* s, int tenant);
/* A borrowed handle to an admitted tenant's instance (owned by the server; never pass it to
 * gz_alg_destroy).  It stays valid until the tenant is released or the server destroyed; every call
 * for one tenant returns the same handle. */
int gz_server_tenant_alg(gz_server* s, int tenant, gz_alg** out);
uint64_t gz_server_reserved_bytes(const gz_server* s);
uint64_t gz_server_budget_bytes(const gz_server* s);

/* Run a mixed trace.  Tenants named in the trace that are not admitted yet are admitted on first
 * use in the order they appear, waiting (queued) until enough admitted tenants finish and release
 * their memory; a tenant is released after its last request when release_when_done != 0.
 * tenant_pairs[t] is tenant t's stream in device memory; tenant_n / tenant_seed give each tenant's
 * parameters for admission (tenant ids index these arrays).  Requests of one tenant are applied in
 * trace order, none before its arrival time; latency = completion - arrival.  Different tenants'
 * requests overlap on the device.  tenant_digest[t] (may be NULL) receives gz_alg_digest of tenant t
 * after its last request; tenant_parents (may be NULL, as may each tenant_parents[t]) receives the
 * tenant's connected_components() parent array (tenant_n[t] DSU roots) after its last request. */
int gz_server_run_trace(gz_server* s, const gz_request* reqs, size_t nreq, const uint32_t* const* tenant_pairs,
                        const uint32_t* tenant_n, const uint64_t* tenant_seed, size_t num_tenants,
                        const gz_alg_config* cfg, int release_when_done, double* latency_s, uint32_t* answers,
                        uint64_t* tenant_digest, uint32_t* const* tenant_parents, gz_serve_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* GZ_SERVING_H */
