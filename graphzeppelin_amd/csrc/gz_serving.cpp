// gz_serving.cpp -- native request loops over device-resident sketches (include/gz_serving.h).
//
// * serve_trace : the batched/served form for one stream.  Update batches are queued on the
//   device without waiting (the instance buffers them and flushes in large launches); every run of
//   consecutive queries is answered against one state with at most one Borůvka pass.
// * plain_trace : the reference-shaped form (per-update pre_insert, host gutters, one launch per full
//   gutter, synchronous queries that flush every gutter first), for comparison on the same GPU.
// * Server      : many streams at once.  Tenants are admitted by device footprint against a memory
//   budget; worker threads pick ready tenants, so a query of one tenant overlaps other tenants'
//   work on their own HIP streams; a tenant's requests stay in trace order.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "../../include/gz_serving.h"
#include "gz_runtime.hpp"

namespace gz {

using Clock = std::chrono::steady_clock;
static double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// Answer the queries reqs[i, j) of one instance against one consistent state.
static void answer_group(GzAlg& A, const gz_request* reqs, size_t i, size_t j, uint32_t* answers,
                         gz_serve_stats& st) {
  bool any_cc = false;
  std::vector<uint32_t> us, vs;
  std::vector<size_t> where;
  for (size_t k = i; k < j; ++k) {
    if (reqs[k].kind == GZ_REQ_CC) {
      any_cc = true;
    } else {
      us.push_back(reqs[k].u);
      vs.push_back(reqs[k].v);
      where.push_back(k);
    }
  }
  uint32_t ncc = 0;
  std::vector<uint8_t> out(us.size());
  const bool cached = A.answer_queries(us.data(), vs.data(), us.size(), any_cc, &ncc, out.data());
  if (cached) ++st.cached_answers;
  else ++st.boruvka_runs;
  ++st.query_groups;
  for (size_t k = i; k < j; ++k)
    if (reqs[k].kind == GZ_REQ_CC && answers) answers[k] = ncc;
  for (size_t q = 0; q < where.size(); ++q)
    if (answers) answers[where[q]] = out[q];
  st.queries += j - i;
}

void serve_trace_body(GzAlg& A, const uint32_t* pairs, bool on_device, const gz_request* reqs, size_t nreq,
                      double* lat, uint32_t* answers, gz_serve_stats& st);

static void check_requests(const GzAlg& A, const gz_request* reqs, size_t nreq) {
  for (size_t i = 0; i < nreq; ++i) {
    const gz_request& r = reqs[i];
    if (r.kind > GZ_REQ_CC) throw GzError(GZ_ERR_INVALID_ARG, "unknown request kind");
    if (r.kind == GZ_REQ_POINT && (r.u >= A.n || r.v >= A.n)) throw GzError(GZ_ERR_INVALID_ARG, "query vertex out of range");
  }
}

// The stream buffer stays valid for the whole call, so device pairs are ingested by reference
// (grouped where they are at the next flush); every reference is released before returning.
void serve_trace(GzAlg& A, const uint32_t* pairs, bool on_device, const gz_request* reqs, size_t nreq, double* lat,
                 uint32_t* answers, gz_serve_stats& st) {
  check_requests(A, reqs, nreq);
  try {
    serve_trace_body(A, pairs, on_device, reqs, nreq, lat, answers, st);
  } catch (...) {
    A.release_stable();
    throw;
  }
}

void serve_trace_body(GzAlg& A, const uint32_t* pairs, bool on_device, const gz_request* reqs, size_t nreq,
                      double* lat, uint32_t* answers, gz_serve_stats& st) {
  st = gz_serve_stats{};
  const auto t0 = Clock::now();
  size_t i = 0;
  while (i < nreq) {
    if (reqs[i].kind == GZ_REQ_UPDATES) {
      const uint32_t* p = pairs + 2 * reqs[i].offset;
      if (on_device) A.ingest_device_stable(p, reqs[i].count);
      else A.ingest_host(p, reqs[i].count, /*stable=*/true);
      st.updates += reqs[i].count;
      if (lat) lat[i] = 0;
      if (answers) answers[i] = 0;
      ++i;
      continue;
    }
    size_t j = i;
    while (j < nreq && reqs[j].kind != GZ_REQ_UPDATES) ++j;
    const auto q0 = Clock::now();
    answer_group(A, reqs, i, j, answers, st);
    const double l = secs(q0, Clock::now());
    for (size_t k = i; k < j; ++k)
      if (lat) lat[k] = l;
    i = j;
  }
  A.flush_all();
  st.wall_s = secs(t0, Clock::now());
}

// Reference-shaped ingest on the host: gutters per vertex, one launch per full gutter.
void plain_trace(GzAlg& A, const uint32_t* pairs, const gz_request* reqs, size_t nreq, double* lat,
                 uint32_t* answers, gz_serve_stats& st) {
  if (A.cfg.batched) throw GzError(GZ_ERR_INVALID_ARG, "plain_trace needs an instance created with batched = 0");
  check_requests(A, reqs, nreq);
  st = gz_serve_stats{};
  const size_t gsize = std::max<size_t>(1, A.desired_updates_per_batch());
  std::vector<std::vector<uint32_t>> gut(A.n);
  auto insert = [&](uint32_t src, uint32_t dst) {
    auto& g = gut[src];
    if (g.capacity() < gsize) g.reserve(gsize);
    g.push_back(dst);
    if (g.size() == gsize) {
      A.apply_update_batch(src, g.data(), g.size());
      g.clear();
    }
  };
  auto flush_gutters = [&] {
    for (uint32_t v = 0; v < A.n; ++v)
      if (!gut[v].empty()) {
        A.apply_update_batch(v, gut[v].data(), gut[v].size());
        gut[v].clear();
      }
  };
  const auto t0 = Clock::now();
  size_t i = 0;
  while (i < nreq) {
    if (reqs[i].kind == GZ_REQ_UPDATES) {
      const uint32_t* p = pairs + 2 * reqs[i].offset;
      for (uint64_t k = 0; k < reqs[i].count; ++k) {
        const uint32_t a = p[2 * k], b = p[2 * k + 1];
        if (a >= A.n || b >= A.n) throw GzError(GZ_ERR_INVALID_ARG, "vertex id out of range");
        A.pre_insert(a, b);
        insert(a, b);
        insert(b, a);
      }
      st.updates += reqs[i].count;
      if (lat) lat[i] = 0;
      if (answers) answers[i] = 0;
      ++i;
      continue;
    }
    size_t j = i;
    while (j < nreq && reqs[j].kind != GZ_REQ_UPDATES) ++j;
    const auto q0 = Clock::now();
    if (!A.has_cached_query(GZ_CONNECTIVITY)) flush_gutters();  // prep_query
    answer_group(A, reqs, i, j, answers, st);
    const double l = secs(q0, Clock::now());
    for (size_t k = i; k < j; ++k)
      if (lat) lat[k] = l;
    i = j;
  }
  flush_gutters();
  A.flush_all();
  st.wall_s = secs(t0, Clock::now());
}

// ---------------------------------------------------------------------------------------------
// Multi-tenant server
// ---------------------------------------------------------------------------------------------
class Server {
 public:
  struct Tenant {
    std::unique_ptr<GzAlg> alg;
    std::unique_ptr<gz_alg> view;  // borrowed handle given out by gz_server_tenant_alg
    uint64_t footprint = 0;
  };
  uint64_t budget;
  uint64_t reserved = 0;
  int workers;
  std::mutex mtx;
  std::map<int, Tenant> tenants;
  int next_id = 0;

  Server(uint64_t b, int w) : budget(b), workers(std::max(1, w)) {}

  int admit(uint32_t n, uint64_t seed, const gz_alg_config& cfg) {
    const uint64_t fp = GzAlg::footprint(n, cfg);
    {
      std::lock_guard<std::mutex> lk(mtx);
      if (reserved + fp > budget)
        throw GzError(GZ_ERR_ADMISSION, "tenant footprint " + std::to_string(fp) + " B exceeds the unreserved budget " +
                                            std::to_string(budget - reserved) + " B");
      reserved += fp;
    }
    std::unique_ptr<GzAlg> a;
    try {
      a.reset(new GzAlg(n, seed, cfg));
    } catch (...) {
      std::lock_guard<std::mutex> lk(mtx);
      reserved -= fp;
      throw;
    }
    std::unique_ptr<gz_alg> view(new gz_alg{a.get()});
    std::lock_guard<std::mutex> lk(mtx);
    const int id = next_id++;
    tenants[id] = Tenant{std::move(a), std::move(view), fp};
    return id;
  }
  void release(int id) {
    std::unique_ptr<GzAlg> a;
    std::unique_ptr<gz_alg> view;
    {
      std::lock_guard<std::mutex> lk(mtx);
      auto it = tenants.find(id);
      if (it == tenants.end()) throw GzError(GZ_ERR_INVALID_ARG, "unknown tenant");
      a = std::move(it->second.alg);
      view = std::move(it->second.view);
      reserved -= it->second.footprint;
      tenants.erase(it);
    }
  }
  GzAlg* get(int id) {
    std::lock_guard<std::mutex> lk(mtx);
    auto it = tenants.find(id);
    if (it == tenants.end()) throw GzError(GZ_ERR_INVALID_ARG, "unknown tenant");
    return it->second.alg.get();
  }
  gz_alg* view(int id) {
    std::lock_guard<std::mutex> lk(mtx);
    auto it = tenants.find(id);
    if (it == tenants.end()) throw GzError(GZ_ERR_INVALID_ARG, "unknown tenant");
    return it->second.view.get();
  }

  // Mixed replay; see gz_server_run_trace.
  void run_trace(const gz_request* reqs, size_t nreq, const uint32_t* const* tpairs, const uint32_t* tn,
                 const uint64_t* tseed, size_t T, const gz_alg_config& cfg, bool release_done, double* lat,
                 uint32_t* answers, uint64_t* digests, uint32_t* const* parents, gz_serve_stats& total);
};

void Server::run_trace(const gz_request* reqs, size_t nreq, const uint32_t* const* tpairs, const uint32_t* tn,
                       const uint64_t* tseed, size_t T, const gz_alg_config& cfg, bool release_done, double* lat,
                       uint32_t* answers, uint64_t* digests, uint32_t* const* parents, gz_serve_stats& total) {
  struct TState {
    std::deque<size_t> todo;  // request indices in trace order
    int id = -1;              // server tenant id once admitted
    bool busy = false, done = false, seen = false;
    uint64_t footprint = 0;
    gz_serve_stats st{};
  };
  std::vector<TState> ts(T);
  std::vector<size_t> admit_order;
  for (size_t i = 0; i < nreq; ++i) {
    const gz_request& r = reqs[i];
    if (r.tenant >= T) throw GzError(GZ_ERR_INVALID_ARG, "request names an unknown tenant");
    if (r.kind > GZ_REQ_CC) throw GzError(GZ_ERR_INVALID_ARG, "unknown request kind");
    if (r.kind == GZ_REQ_POINT && (r.u >= tn[r.tenant] || r.v >= tn[r.tenant]))
      throw GzError(GZ_ERR_INVALID_ARG, "query vertex out of range");
    ts[r.tenant].todo.push_back(i);
    if (!ts[r.tenant].seen) {
      ts[r.tenant].seen = true;
      admit_order.push_back(r.tenant);
    }
  }
  for (size_t t : admit_order) {
    ts[t].footprint = GzAlg::footprint(tn[t], cfg);
    if (ts[t].footprint > budget) throw GzError(GZ_ERR_ADMISSION, "a tenant can never fit the memory budget");
  }
  total = gz_serve_stats{};
  const auto t0 = Clock::now();
  std::mutex m;
  std::condition_variable cv;
  size_t next_admit = 0, finished = 0;
  std::string err;
  int err_code = 0;

  auto now_s = [&] { return secs(t0, Clock::now()); };

  // Admit waiting tenants (first fit, in order of first appearance).  Caller holds m.
  auto admit_some = [&] {
    while (next_admit < admit_order.size()) {
      bool any = false;
      for (size_t k = next_admit; k < admit_order.size(); ++k) {
        const size_t t = admit_order[k];
        if (ts[t].id >= 0 || ts[t].done) continue;
        uint64_t free_b;
        {
          std::lock_guard<std::mutex> lk(mtx);
          free_b = budget - reserved;
        }
        if (ts[t].footprint > free_b) continue;
        ts[t].id = admit(tn[t], tseed[t], cfg);
        any = true;
      }
      while (next_admit < admit_order.size() && ts[admit_order[next_admit]].id >= 0) ++next_admit;
      if (!any) break;
    }
  };

  auto worker = [&] {
    std::unique_lock<std::mutex> lk(m);
    while (true) {
      if (!err.empty() || finished == admit_order.size()) return;
      // pick the admitted, idle tenant whose next request arrived first (first come, first served
      // across tenants, so no tenant waits behind others that keep arriving); else wait
      size_t pick = T;
      double next_arrival = 1e300, oldest = 1e300;
      const double now = now_s();
      for (size_t t : admit_order) {
        TState& s = ts[t];
        if (s.id < 0 || s.busy || s.done || s.todo.empty()) continue;
        const double arr = reqs[s.todo.front()].arrival_s;
        if (arr <= now) {
          if (arr < oldest) {
            oldest = arr;
            pick = t;
          }
        } else {
          next_arrival = std::min(next_arrival, arr);
        }
      }
      if (pick == T) {
        size_t active = 0;
        for (size_t t : admit_order)
          if (ts[t].id >= 0 && !ts[t].done) ++active;
        if (active == 0 && next_admit < admit_order.size()) {
          admit_some();
          bool admitted = false;
          for (size_t t : admit_order)
            if (ts[t].id >= 0 && !ts[t].done) admitted = true;
          if (!admitted) {
            err = "waiting tenants can never be admitted: nothing releases memory (release_when_done = 0)";
            err_code = GZ_ERR_ADMISSION;
            cv.notify_all();
            return;
          }
          continue;
        }
        if (next_arrival < 1e299)
          cv.wait_for(lk, std::chrono::duration<double>(std::max(0.0, next_arrival - now_s())));
        else
          cv.wait_for(lk, std::chrono::milliseconds(5));
        continue;
      }
      TState& s = ts[pick];
      s.busy = true;
      GzAlg* A = get(s.id);
      // take every request of this tenant that has arrived, up to and including one query group
      std::vector<size_t> mine;
      while (!s.todo.empty() && reqs[s.todo.front()].arrival_s <= now_s()) {
        const size_t i = s.todo.front();
        if (reqs[i].kind != GZ_REQ_UPDATES && !mine.empty() && reqs[mine.back()].kind == GZ_REQ_UPDATES) break;
        mine.push_back(i);
        s.todo.pop_front();
        if (reqs[i].kind != GZ_REQ_UPDATES &&
            (s.todo.empty() || reqs[s.todo.front()].kind == GZ_REQ_UPDATES))
          break;
      }
      lk.unlock();
      try {
        size_t a = 0;
        while (a < mine.size()) {
          const gz_request& r = reqs[mine[a]];
          if (r.kind == GZ_REQ_UPDATES) {
            A->ingest_device(tpairs[pick] + 2 * r.offset, r.count);
            s.st.updates += r.count;
            if (lat) lat[mine[a]] = now_s() - r.arrival_s;
            if (answers) answers[mine[a]] = 0;
            ++a;
            continue;
          }
          size_t b = a;
          while (b < mine.size() && reqs[mine[b]].kind != GZ_REQ_UPDATES) ++b;
          // the group is contiguous in this tenant's order; answer it with a local request copy
          std::vector<gz_request> grp;
          for (size_t k = a; k < b; ++k) grp.push_back(reqs[mine[k]]);
          std::vector<uint32_t> ans(grp.size());
          answer_group(*A, grp.data(), 0, grp.size(), ans.data(), s.st);
          const double done_t = now_s();
          for (size_t k = a; k < b; ++k) {
            if (answers) answers[mine[k]] = ans[k - a];
            if (lat) lat[mine[k]] = done_t - reqs[mine[k]].arrival_s;
          }
          a = b;
        }
        bool last = false;
        {
          std::lock_guard<std::mutex> g(m);
          last = s.todo.empty();
        }
        if (last) {
          A->flush_all();
          if (digests) digests[pick] = A->digest();
          if (parents && parents[pick]) A->connected_components(parents[pick]);
        }
        lk.lock();
        s.busy = false;
        if (last) {
          s.done = true;
          ++finished;
          if (release_done) release(s.id);
          admit_some();
        }
      } catch (const GzError& e) {
        lk.lock();
        if (err.empty()) { err = e.what(); err_code = e.code; }
        cv.notify_all();
        return;
      } catch (const std::exception& e) {
        lk.lock();
        if (err.empty()) { err = e.what(); err_code = GZ_ERR_DEVICE; }
        cv.notify_all();
        return;
      }
      cv.notify_all();
    }
  };

  {
    std::lock_guard<std::mutex> g(m);
    admit_some();
  }
  std::vector<std::thread> th;
  for (int w = 0; w < workers; ++w) th.emplace_back(worker);
  for (auto& x : th) x.join();
  if (!err.empty()) throw GzError(err_code, err);
  total.wall_s = now_s();
  for (auto& s : ts) {
    total.updates += s.st.updates;
    total.queries += s.st.queries;
    total.boruvka_runs += s.st.boruvka_runs;
    total.cached_answers += s.st.cached_answers;
    total.query_groups += s.st.query_groups;
  }
}

}  // namespace gz

// ================================================================================================
// C ABI
// ================================================================================================
struct gz_server {
  gz::Server* s;
};

extern "C" {

int gz_alg_serve_trace(gz_alg* a, const uint32_t* pairs, int on_device, const gz_request* reqs, size_t nreq,
                       double* lat, uint32_t* answers, gz_serve_stats* stats) {
  return gz::guarded_call([&] {
    if (!a || (!pairs && nreq) || (!reqs && nreq)) throw gz::GzError(GZ_ERR_INVALID_ARG, "null argument");
    gz_serve_stats st{};
    gz::serve_trace(*a->a, pairs, on_device != 0, reqs, nreq, lat, answers, st);
    if (stats) *stats = st;
  });
}

int gz_alg_plain_trace(gz_alg* a, const uint32_t* pairs, const gz_request* reqs, size_t nreq, double* lat,
                       uint32_t* answers, gz_serve_stats* stats) {
  return gz::guarded_call([&] {
    if (!a || (!pairs && nreq) || (!reqs && nreq)) throw gz::GzError(GZ_ERR_INVALID_ARG, "null argument");
    gz_serve_stats st{};
    gz::plain_trace(*a->a, pairs, reqs, nreq, lat, answers, st);
    if (stats) *stats = st;
  });
}

int gz_server_create(uint64_t budget, int workers, gz_server** out) {
  return gz::guarded_call([&] {
    if (!out) throw gz::GzError(GZ_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    gz::require_device();
    if (budget == 0) {
      size_t fr = 0, tot = 0;
      GZ_HIP(hipMemGetInfo(&fr, &tot));
      budget = fr;
    }
    *out = new gz_server{new gz::Server(budget, workers)};
  });
}
int gz_server_destroy(gz_server* s) {
  return gz::guarded_call([&] {
    if (!s) return;
    delete s->s;
    delete s;
  });
}
int gz_server_admit(gz_server* s, uint32_t n, uint64_t seed, const gz_alg_config* cfg, int* tenant) {
  if (tenant) *tenant = -1;
  return gz::guarded_call([&] {
    if (!s || !tenant) throw gz::GzError(GZ_ERR_INVALID_ARG, "null argument");
    gz_alg_config c;
    gz_alg_config_default(&c);
    if (cfg) c = *cfg;
    *tenant = s->s->admit(n, seed, c);
  });
}
int gz_server_release(gz_server* s, int tenant) {
  return gz::guarded_call([&] {
    if (!s) throw gz::GzError(GZ_ERR_INVALID_ARG, "null server");
    s->s->release(tenant);
  });
}
int gz_server_tenant_alg(gz_server* s, int tenant, gz_alg** out) {
  return gz::guarded_call([&] {
    if (!s || !out) throw gz::GzError(GZ_ERR_INVALID_ARG, "null argument");
    // borrowed handle: the server owns the instance and the handle, which stays valid until the
    // tenant is released; gz_alg_destroy must not be called on it
    *out = s->s->view(tenant);
  });
}
uint64_t gz_server_reserved_bytes(const gz_server* s) {
  if (!s) return 0;
  std::lock_guard<std::mutex> lk(s->s->mtx);
  return s->s->reserved;
}
uint64_t gz_server_budget_bytes(const gz_server* s) { return s ? s->s->budget : 0; }

int gz_server_run_trace(gz_server* s, const gz_request* reqs, size_t nreq, const uint32_t* const* tpairs,
                        const uint32_t* tn, const uint64_t* tseed, size_t T, const gz_alg_config* cfg,
                        int release_done, double* lat, uint32_t* answers, uint64_t* digests,
                        uint32_t* const* parents, gz_serve_stats* stats) {
  return gz::guarded_call([&] {
    if (!s || (!reqs && nreq) || !tpairs || !tn || !tseed) throw gz::GzError(GZ_ERR_INVALID_ARG, "null argument");
    gz_alg_config c;
    gz_alg_config_default(&c);
    if (cfg) c = *cfg;
    gz_serve_stats st{};
    s->s->run_trace(reqs, nreq, tpairs, tn, tseed, T, c, release_done != 0, lat, answers, digests, parents, st);
    if (stats) *stats = st;
  });
}

}  // extern "C"
