// Continuous-batching request loop over a loaded index (no reference
// counterpart; it replaces the synchronous batching of the CLI's `load`
// command, tools/katoann.cpp:267-337, whose latency percentile rule
// sorted[min(n-1, ceil(q n)-1)] (tools/katoann.cpp:325-327) it keeps).
//
// Requests queue in arrival order. Each of max_inflight worker threads takes
// the oldest request and the later queued requests with the same search
// parameters (up to max_batch queries), reserves the batch's search footprint
// against the HBM budget, and runs it through katoann_search_ex on a search
// context of its own (its own HIP stream), so batches of different workers
// overlap on the GPU. Host staging buffers are pinned (hipHostMalloc) so the
// copies in and out are asynchronous DMA. The budget bounds the sum of the
// footprints of the batches in flight: a search may not grow its buffers past
// its batch's reservation, and a survivor-pool overflow whose rerun needs more
// room reserves that room first (new batches wait meanwhile). The index keeps
// an idle search context's buffers for reuse, so the device memory held for
// searches is at most max_inflight times the largest batch footprint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/katoann_c.h"

extern std::string &katoann_error_slot();

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  uint32_t nq = 0;
  katoann_search_params p{};
  bool has_model = false;
  katoann_nprobe_model model{};
  int32_t normalize = 0;
  std::vector<float> queries;
  std::vector<uint32_t> ids;
  std::vector<float> dists;
  Clock::time_point arrival, done_at;
  bool done = false;
  int status = KATOANN_OK;
  std::string error;
};

bool same_params(const Request &a, const Request &b) {
  if (a.p.k != b.p.k || a.p.nprobe != b.p.nprobe || a.p.ranking != b.p.ranking || a.p.survivors != b.p.survivors ||
      a.p.selection != b.p.selection || a.normalize != b.normalize || a.has_model != b.has_model)
    return false;
  return !a.has_model || std::memcmp(&a.model, &b.model, sizeof(a.model)) == 0;
}

// Pinned host buffer that grows on demand.
struct Pinned {
  void *p = nullptr;
  size_t bytes = 0;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  void *ensure(size_t n) {
    if (n > bytes || !p) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      bytes = 0;
      if (hipHostMalloc(&p, std::max<size_t>(n, 4096), hipHostMallocDefault) != hipSuccess)
        throw std::runtime_error("cannot allocate pinned staging memory");
      bytes = std::max<size_t>(n, 4096);
    }
    return p;
  }
};

}  // namespace

struct katoann_server {
  katoann_index *ix = nullptr;
  uint32_t dim = 0;
  katoann_server_options opt{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  uint64_t reserved = 0, peak_reserved = 0, peak_used = 0;
  uint32_t rerun_waiting = 0;  // batches waiting for room to run again after a pool overflow
  uint32_t inflight = 0, peak_inflight = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  // statistics of completed requests
  std::vector<double> lat_ms;
  uint64_t n_queries = 0, n_batches = 0, batch_queries = 0;
  bool any_arrival = false;
  Clock::time_point first_arrival, last_done;

  uint64_t footprint(const Request &r, uint32_t nq) const {
    katoann_search_extras x{};
    x.normalize = r.normalize;
    x.nprobe_model = r.has_model ? &r.model : nullptr;
    uint64_t b = 0;
    if (katoann_search_footprint(ix, nq, &r.p, &x, &b) != KATOANN_OK)
      throw std::invalid_argument(katoann_error_slot());
    return b;
  }

  // Under mu: the next batch that fits, or an empty vector.
  std::vector<Request *> take(uint64_t &bytes) {
    std::vector<Request *> batch;
    if (rerun_waiting) return batch;  // a batch in flight waits for room to run again: it goes first
    Request *head = nullptr;
    uint64_t need = 0;
    while (!queue.empty()) {
      head = queue.front();
      try {
        need = footprint(*head, head->nq);
        break;
      } catch (const std::exception &e) {  // checked at submission; fail it rather than the worker
        head->status = KATOANN_INVALID_ARGUMENT;
        head->error = e.what();
        head->done = true;
        head->done_at = Clock::now();
        queue.pop_front();
        head = nullptr;
        cv_done.notify_all();
      }
    }
    if (!head) return batch;
    if (reserved + need > opt.hbm_budget) return batch;  // wait for memory: no request overtakes the oldest
    batch.push_back(head);
    uint32_t nq = head->nq;
    for (size_t i = 1; i < queue.size(); ++i) {
      Request *r = queue[i];
      if (!same_params(*head, *r) || nq + r->nq > opt.max_batch) continue;
      uint64_t n2 = 0;
      try {
        n2 = footprint(*head, nq + r->nq);
      } catch (const std::exception &) {
        break;  // the merged shape is not servable: run what is gathered
      }
      if (reserved + n2 > opt.hbm_budget) break;
      batch.push_back(r);
      nq += r->nq;
      need = n2;
    }
    for (Request *r : batch) queue.erase(std::find(queue.begin(), queue.end(), r));
    bytes = need;
    return batch;
  }

  void worker() {
    Pinned qbuf, ibuf, dbuf;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      std::vector<Request *> batch;
      uint64_t bytes = 0;
      cv_work.wait(lk, [&] {
        if (stop && queue.empty()) return true;
        batch = take(bytes);
        return !batch.empty();
      });
      if (batch.empty()) return;  // stopping, queue drained
      reserved += bytes;
      peak_reserved = std::max(peak_reserved, reserved);
      ++inflight;
      peak_inflight = std::max(peak_inflight, inflight);
      lk.unlock();

      const Request &h = *batch.front();
      uint32_t nq = 0;
      for (Request *r : batch) nq += r->nq;
      int status = KATOANN_OK;
      std::string error;
      uint64_t used = 0;
      try {
        float *q = static_cast<float *>(qbuf.ensure(static_cast<size_t>(nq) * dim * 4));
        uint32_t *ids = static_cast<uint32_t *>(ibuf.ensure(static_cast<size_t>(nq) * h.p.k * 4));
        float *dists = static_cast<float *>(dbuf.ensure(static_cast<size_t>(nq) * h.p.k * 4));
        size_t off = 0;
        for (Request *r : batch) {
          std::memcpy(q + off, r->queries.data(), r->queries.size() * 4);
          off += r->queries.size();
        }
        for (;;) {
          // the search may not grow its buffers past the reservation: a pool
          // overflow whose rerun needs more room returns, and the batch
          // reserves that room (ahead of new batches) before running again
          katoann_search_extras x{};
          x.normalize = h.normalize;
          x.nprobe_model = h.has_model ? &h.model : nullptr;
          x.max_bytes = bytes;
          uint64_t need = 0;
          x.needed_bytes = &need;
          status = katoann_search_ex(ix, q, nq, &h.p, &x, ids, dists);
          if (status != KATOANN_RUNTIME_ERROR || need <= bytes) {
            used = need;
            break;
          }
          lk.lock();
          reserved -= bytes;
          bytes = 0;
          if (need > opt.hbm_budget) {
            lk.unlock();
            katoann_error_slot() = "survivor pool overflow: the batch needs " + std::to_string(need) +
                                   " bytes of HBM, above the budget of " + std::to_string(opt.hbm_budget);
            break;
          }
          ++rerun_waiting;
          cv_done.wait(lk, [&] { return reserved + need <= opt.hbm_budget; });
          --rerun_waiting;
          reserved += need;
          peak_reserved = std::max(peak_reserved, reserved);
          bytes = need;
          lk.unlock();
        }
        if (status != KATOANN_OK) {
          error = katoann_error_slot();
        } else {
          size_t o = 0;
          for (Request *r : batch) {
            const size_t m = static_cast<size_t>(r->nq) * h.p.k;
            r->ids.assign(ids + o, ids + o + m);
            r->dists.assign(dists + o, dists + o + m);
            o += m;
          }
        }
      } catch (const std::exception &e) {
        status = KATOANN_RUNTIME_ERROR;
        error = e.what();
      }
      const auto now = Clock::now();

      lk.lock();
      reserved -= bytes;
      peak_used = std::max(peak_used, used);
      --inflight;
      ++n_batches;
      batch_queries += nq;
      for (Request *r : batch) {
        r->status = status;
        r->error = error;
        r->done = true;
        r->done_at = now;
        r->queries.clear();
        r->queries.shrink_to_fit();
        if (status == KATOANN_OK) {
          lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->arrival).count());
          n_queries += r->nq;
        }
      }
      last_done = now;
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

namespace {

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  const size_t i = static_cast<size_t>(std::ceil(q * static_cast<double>(n)));
  return v[std::min(n - 1, i == 0 ? 0 : i - 1)];
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return KATOANN_OK;
  } catch (const std::invalid_argument &e) {
    katoann_error_slot() = e.what();
    return KATOANN_INVALID_ARGUMENT;
  } catch (const std::logic_error &e) {
    katoann_error_slot() = e.what();
    return KATOANN_LOGIC_ERROR;
  } catch (const std::exception &e) {
    katoann_error_slot() = e.what();
    return KATOANN_RUNTIME_ERROR;
  }
}

}  // namespace

extern "C" {

void katoann_default_server_options(katoann_server_options *o) { *o = katoann_server_options{4u, 4096u, 0ull}; }

int katoann_server_create(katoann_index *h, const katoann_server_options *o, katoann_server **out) {
  return guarded([&] {
    if (!h || !out) throw std::invalid_argument("null argument");
    if (katoann_nlist(h) == 0) throw std::logic_error("server on an index that is not loaded");
    auto s = std::make_unique<katoann_server>();
    s->ix = h;
    s->dim = katoann_dim(h);
    katoann_default_server_options(&s->opt);
    if (o) s->opt = *o;
    if (s->opt.max_inflight == 0 || s->opt.max_inflight > 64)
      throw std::invalid_argument("max_inflight must be in [1, 64]");
    if (s->opt.max_batch == 0) throw std::invalid_argument("max_batch must be >= 1");
    if (s->opt.hbm_budget == 0) {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw std::runtime_error("hipMemGetInfo failed");
      s->opt.hbm_budget = static_cast<uint64_t>(static_cast<double>(free_b) * 0.9);
    }
    for (uint32_t i = 0; i < s->opt.max_inflight; ++i) s->workers.emplace_back([p = s.get()] { p->worker(); });
    *out = s.release();
  });
}

void katoann_server_destroy(katoann_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  delete s;
}

int katoann_server_submit(katoann_server *s, const float *queries, uint32_t nq, const katoann_search_params *p,
                          const katoann_nprobe_model *model, int32_t normalize, uint64_t *ticket) {
  return guarded([&] {
    if (!s || !p || !ticket || (nq && !queries)) throw std::invalid_argument("null argument");
    if (nq == 0) throw std::invalid_argument("a request needs at least one query");
    auto r = std::make_unique<Request>();
    r->nq = nq;
    r->p = *p;
    r->has_model = model != nullptr;
    if (model) r->model = *model;
    r->normalize = normalize;
    if (nq > s->opt.max_batch) throw std::invalid_argument("request larger than the server's max_batch");
    const uint64_t need = s->footprint(*r, nq);  // also checks the parameters
    if (need > s->opt.hbm_budget)
      throw std::invalid_argument("request needs " + std::to_string(need) + " bytes of HBM, above the budget of " +
                                  std::to_string(s->opt.hbm_budget));
    r->queries.assign(queries, queries + static_cast<size_t>(nq) * s->dim);
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) throw std::logic_error("server is stopping");
    r->arrival = Clock::now();
    if (!s->any_arrival) {
      s->any_arrival = true;
      s->first_arrival = r->arrival;
    }
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->queue.push_back(r.get());
    s->reqs.emplace(r->ticket, std::move(r));
    s->cv_work.notify_one();
  });
}

int katoann_server_poll(katoann_server *s, uint64_t ticket, int32_t *done) {
  return guarded([&] {
    if (!s || !done) throw std::invalid_argument("null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) throw std::invalid_argument("unknown ticket");
    *done = it->second->done ? 1 : 0;
  });
}

int katoann_server_wait(katoann_server *s, uint64_t ticket, uint32_t *ids, float *dists, double *latency_ms) {
  std::unique_ptr<Request> r;
  const int rc = guarded([&] {
    if (!s) throw std::invalid_argument("null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) throw std::invalid_argument("unknown ticket");
    Request *raw = it->second.get();
    s->cv_done.wait(lk, [&] { return raw->done; });
    r = std::move(it->second);
    s->reqs.erase(it);
  });
  if (rc != KATOANN_OK) return rc;
  if (r->status != KATOANN_OK) {
    katoann_error_slot() = r->error;
    return r->status;
  }
  if (ids) std::memcpy(ids, r->ids.data(), r->ids.size() * 4);
  if (dists) std::memcpy(dists, r->dists.data(), r->dists.size() * 4);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->done_at - r->arrival).count();
  return KATOANN_OK;
}

int katoann_server_drain(katoann_server *s) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->inflight == 0; });
  });
}

int katoann_server_stats_get(katoann_server *s, katoann_server_stats *out) {
  return guarded([&] {
    if (!s || !out) throw std::invalid_argument("null argument");
    std::lock_guard<std::mutex> g(s->mu);
    katoann_server_stats st{};
    st.requests = s->lat_ms.size();
    st.queries = s->n_queries;
    st.batches = s->n_batches;
    st.p50_ms = percentile(s->lat_ms, 0.50);
    st.p95_ms = percentile(s->lat_ms, 0.95);
    st.p99_ms = percentile(s->lat_ms, 0.99);
    st.max_ms = s->lat_ms.empty() ? 0.0 : *std::max_element(s->lat_ms.begin(), s->lat_ms.end());
    st.mean_batch = s->n_batches ? static_cast<double>(s->batch_queries) / static_cast<double>(s->n_batches) : 0.0;
    const double span =
        s->any_arrival && st.requests ? std::chrono::duration<double>(s->last_done - s->first_arrival).count() : 0.0;
    st.achieved_qps = span > 0 ? static_cast<double>(s->n_queries) / span : 0.0;
    st.budget = s->opt.hbm_budget;
    st.peak_reserved = s->peak_reserved;
    st.peak_used = s->peak_used;
    st.peak_inflight = s->peak_inflight;
    *out = st;
  });
}

}  // extern "C"
