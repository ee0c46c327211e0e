// serve.hip — continuous-batching request loop over device-resident graphs.
// No reference counterpart (the reference runs one query per process).
//
// A scheduler thread owns the context's stream while the server runs:
//  * admission: a submitted request is admitted when the number of requests
//    on the device is below max_inflight and the HBM its intermediates need
//    (per-graph auxiliary arrays it would create plus its share of the list
//    scratch) fits the budget; a request that could never fit is rejected
//    at submit with JGPM_E_BUDGET;
//  * slicing: every admitted request is cut into root slices of bounded work
//    (edges of the slice times a per-pattern cost weight), so expensive
//    queries cannot hold the device for long;
//  * iteration: each loop takes the next slice of every admitted request and
//    runs them all in one persistent batched launch; requests that arrived
//    meanwhile are admitted before the next launch, so their work overlaps
//    with the tails of earlier ones on the device;
//  * heavy requests: a request whose work (edges of its range times the
//    pattern weight) exceeds one slice would fill the device by itself, so
//    it is not sliced: it runs whole with the single-query kernels (its count
//    is then exactly jgpm_count's, from the same kernels, including the
//    degree-order relabelling a whole-graph 4-cycle count uses).  Heavy
//    requests run on lanes: each lane is a thread with a context of its own
//    (stream, accumulators, counters, scratch) on the same device, so up to
//    heavy_lanes heavy requests and the light batches overlap on the device,
//    filling the tail of one query's launches with the next query's work.
//    Before a heavy request goes to a lane, the scheduler builds every
//    per-graph array its kernels would create on first use (on the graph's
//    own context, waited for), so a lane only reads the shared graph;
//  * completion: a request finishes when its last slice is counted; its
//    latency is submit -> result visible on the host.
// Counts are sums over disjoint root slices, so a result is bit-identical to
// the single-query count of the same (graph, pattern, root range).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "jgpm_device.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct Req {
  uint64_t ticket;
  jgpm_graph_t g;
  int pattern;
  int64_t rb, re, cur;
  uint64_t count = 0;
  int64_t footprint = 0;
  bool heavy = false;       // runs whole, with the single-query kernels (not sliced into batches)
  bool dispatched = false;  // a heavy request handed to a lane
  int status = JGPM_OK;
  clk::time_point t_submit;
};

struct HeavyDone {
  uint64_t ticket;
  uint64_t count;
  int status;
};

// One heavy request at a time, on a context of its own.
struct Lane {
  jgpm_ctx_s* ctx = nullptr;
  std::thread th;
  bool busy = false;  // guarded by the server's mutex, as is job
  Req job;
};

struct Done {
  uint64_t count;
  double latency_us;
  int status;
};

// relative cost per edge of a slice, used only to bound slice work
double pattern_weight(int p) {
  switch (p) {
    case JGPM_TC: case JGPM_P2: return 1;
    case JGPM_P13: case JGPM_P3: case JGPM_P5: case JGPM_P7: case JGPM_P9: return 4;
    case JGPM_P12: case JGPM_P11: return 16;
    case JGPM_P14: case JGPM_P1: case JGPM_P4: case JGPM_P8: return 32;
    case JGPM_P6: case JGPM_P10: case JGPM_P15: return 128;
    case JGPM_P16: return 512;
  }
  return 8;
}

}  // namespace

struct jgpm_serve_s {
  jgpm_ctx_s* ctx = nullptr;
  jgpm_serve_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req> pending;
  std::vector<Req> active;
  std::unordered_map<uint64_t, Done> done;  // finished, not yet retrieved
  std::unordered_set<uint64_t> live;         // submitted, not yet retrieved
  uint64_t next_ticket = 1;
  int64_t inflight_bytes = 0;
  bool stop = false;
  int fatal = JGPM_OK;
  std::thread th;
  int64_t slice_edges = 1 << 22;
  std::vector<Lane*> lanes;
  std::condition_variable cv_lane;
  std::vector<HeavyDone> heavy_done;  // finished on a lane, not yet merged by the scheduler
  bool lanes_stop = false;
};

static bool is_dag_pattern(int p) { return jgpm::clique_k(p) != 0 && p != JGPM_I_K4; }

// HBM the request's intermediates need beyond the resident graph.
static int64_t footprint(jgpm_graph_s* g, int p, const jgpm_ctx_s* c, int grid) {
  int64_t b = 64;  // count slot and queue counters
  if (!is_dag_pattern(p)) {
    if (!g->d_src) b += 4 * g->sym.nnz;
    if (!g->d_tri && jgpm::query_needs_tri(c, p)) b += 4 * g->sym.nnz;
    if (jk::sym_heavy_pattern(p))  // the heavy-edge lists of the context that runs it
      b += 12 * std::max<int64_t>(0, std::min<int64_t>(g->sym.nnz, (int64_t)1 << 26) - c->hcap);
    if (!g->d_nsum && p == JGPM_P3) b += 8 * (int64_t)g->sym.nv;
    b += (int64_t)grid * 4 * 5 * 4 * std::max(1, g->sym_mmin);  // list scratch of the batch kernel
  } else if (g->dag.max_deg > jk::tier_max_deg(jk::TIER_LDS16)) {
    // the 2048-wide tier's row scratch: slices of the fused kernel, or for
    // 5- to 7-cliques the 512 KiB row slots of the split one (<= 1 GiB)
    const int64_t fused = (int64_t)4 * c->num_cu * (int64_t)jk::tier_scratch_words(jk::TIER_GLB32) * 8;
    b += jgpm::clique_k(p) >= 5 ? std::max<int64_t>(fused, (int64_t)1 << 30) : fused;
  }
  (void)c;
  return b;
}

// Runs a heavy request whole on lane context lc; the graph's lazy arrays are
// built (prepare_pattern), so the enqueue only reads the graph, under its
// context's lock like every other reader.
static int run_heavy(jgpm_ctx_s* main, jgpm_ctx_s* lc, const Req& r, uint64_t* count) {
  std::lock_guard<std::mutex> own(lc->mu);
  int st = hipSetDevice(lc->dev) == hipSuccess ? JGPM_OK : JGPM_E_HIP;
  {
    jgpm::ExecScope scope(lc);
    std::lock_guard<std::mutex> gl(main->mu);
    if (!st) st = jgpm::ctx_reset_counters(lc, lc->st);
    if (!st) {
      lc->nprof = 0;
      st = jgpm::enqueue_pattern(r.g, r.pattern, r.rb, r.re, 0, lc->st);
    }
  }
  uint64_t hc[CTX_SLOTS];
  if (!st && hipMemcpyAsync(hc, lc->d_out, sizeof(hc), hipMemcpyDeviceToHost, lc->st) != hipSuccess) st = JGPM_E_HIP;
  if (hipStreamSynchronize(lc->st) != hipSuccess && !st) st = JGPM_E_HIP;
  if (!st) *count = jgpm::finish_pattern(r.pattern, hc);
  return st;
}

static void lane_loop(jgpm_serve_s* s, Lane* L) {
  for (;;) {
    Req r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_lane.wait(lk, [&] { return s->lanes_stop || L->busy; });
      if (!L->busy) return;
      r = L->job;
    }
    uint64_t count = 0;
    const int st = run_heavy(s->ctx, L->ctx, r, &count);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      L->busy = false;
      s->heavy_done.push_back(HeavyDone{r.ticket, count, st});
    }
    s->cv_work.notify_all();
  }
}

static Lane* free_lane(jgpm_serve_s* s) {
  for (Lane* L : s->lanes)
    if (!L->busy) return L;
  return nullptr;
}

static void serve_loop(jgpm_serve_s* s) {
  jgpm_ctx_s* c = s->ctx;
  std::vector<jgpm_query> qs;
  std::vector<size_t> owner;
  std::vector<uint64_t> h;
  for (;;) {
    std::vector<size_t> to_lane;  // heavy requests to hand to lanes, oldest first
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // work: an admissible arrival, lane results, light slices, or a heavy
      // request and a free lane
      auto has_work = [&] {
        if (!s->heavy_done.empty()) return true;
        if (!s->pending.empty() && (int)s->active.size() < s->cfg.max_inflight &&
            (s->active.empty() || s->inflight_bytes + s->pending.front().footprint <= s->cfg.hbm_budget_bytes))
          return true;
        for (const Req& r : s->active) {
          if (r.cur >= r.re) continue;
          if (!r.heavy || (!r.dispatched && free_lane(s))) return true;
        }
        return false;
      };
      s->cv_work.wait(lk, [&] { return has_work() || (s->stop && s->active.empty() && s->pending.empty()); });
      if (s->stop && s->active.empty() && s->pending.empty()) return;
      // admission in arrival order
      while (!s->pending.empty() && (int)s->active.size() < s->cfg.max_inflight) {
        Req& r = s->pending.front();
        if (!s->active.empty() && s->inflight_bytes + r.footprint > s->cfg.hbm_budget_bytes) break;
        s->inflight_bytes += r.footprint;
        s->active.push_back(r);
        s->pending.pop_front();
      }
      // results of the lanes
      for (const HeavyDone& d : s->heavy_done)
        for (Req& r : s->active)
          if (r.ticket == d.ticket) {
            r.count = d.count;
            r.status = d.status;
            r.cur = r.re;
          }
      s->heavy_done.clear();
      int free = 0;
      for (Lane* L : s->lanes) free += !L->busy;
      for (size_t i = 0; i < s->active.size() && (int)to_lane.size() < free; ++i) {
        const Req& r = s->active[i];
        if (r.heavy && !r.dispatched && r.cur < r.re) to_lane.push_back(i);
      }
    }
    // heavy requests to free lanes (the active list is only touched by this thread)
    for (size_t i : to_lane) {
      Req& r = s->active[i];
      const int rc = jgpm::prepare_pattern(r.g, r.pattern, r.rb, r.re);
      std::lock_guard<std::mutex> lk(s->mu);
      r.dispatched = true;
      if (rc) {
        r.status = rc;
        r.cur = r.re;
        continue;
      }
      Lane* L = free_lane(s);  // lanes only turn free meanwhile, so one is
      L->job = r;
      L->busy = true;
      s->cv_lane.notify_all();
    }
    // next slice of every light active request
    qs.clear();
    owner.clear();
    for (size_t i = 0; i < s->active.size(); ++i) {
      Req& r = s->active[i];
      if (r.cur >= r.re || r.heavy) continue;
      jgpm_graph_s* g = r.g;
      const std::vector<int64_t>& rp = is_dag_pattern(r.pattern) ? g->dag_rp_h : g->sym_rp_h;
      const int64_t budget = std::max<int64_t>(1, (int64_t)(s->slice_edges / pattern_weight(r.pattern)));
      int64_t end = std::upper_bound(rp.begin() + r.cur, rp.begin() + r.re + 1, rp[r.cur] + budget) - rp.begin() - 1;
      end = std::max(end, r.cur + 1);
      if (s->cfg.slice_roots > 0) end = std::min(end, r.cur + s->cfg.slice_roots);
      end = std::min(end, r.re);
      qs.push_back(jgpm_query{g, r.pattern, 0, r.cur, end});
      owner.push_back(i);
      r.cur = end;
    }
    int st = JGPM_OK;
    if (!qs.empty()) {
      {
        // the context lock covers the enqueue; later work on the stream runs after these launches
        std::lock_guard<std::mutex> lk(c->mu);
        st = (hipSetDevice(c->dev) == hipSuccess) ? JGPM_OK : JGPM_E_HIP;
        if (!st) st = jgpm::ctx_reset_counters(c, c->st);
        if (!st) {
          c->nprof = 0;
          st = jgpm::enqueue_batch_q(c, qs.data(), (int)qs.size(), c->st);
        }
        h.assign(qs.size(), 0);
        if (!st && hipMemcpyAsync(h.data(), c->d_bcnt, sizeof(uint64_t) * qs.size(), hipMemcpyDeviceToHost, c->st) !=
                       hipSuccess)
          st = JGPM_E_HIP;
      }
      if (hipStreamSynchronize(c->st) != hipSuccess && !st) st = JGPM_E_HIP;
    }
    const auto now = clk::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t k = 0; k < qs.size(); ++k) s->active[owner[k]].count += h[k];
    // finished requests (all light ones after a failed batch) leave the device;
    // a heavy request on a lane finishes when its result is merged
    std::vector<Req> keep;
    for (Req& r : s->active) {
      const bool failed = st && !r.heavy;
      if (failed || r.cur >= r.re) {
        double us = std::chrono::duration<double, std::micro>(now - r.t_submit).count();
        s->done[r.ticket] = Done{r.count, us, failed ? st : r.status};
        s->inflight_bytes -= r.footprint;
        if (r.status == JGPM_E_HIP && !s->fatal) s->fatal = r.status;
      } else {
        keep.push_back(r);
      }
    }
    s->active.swap(keep);
    if (st) s->fatal = st;
    s->cv_done.notify_all();
  }
}

// Ends the lane threads (idle once the scheduler has drained) and their contexts.
static void stop_lanes(jgpm_serve_s* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->lanes_stop = true;
  }
  s->cv_lane.notify_all();
  for (Lane* L : s->lanes) {
    if (L->th.joinable()) L->th.join();
    if (L->ctx) (void)jgpm_shutdown(L->ctx);
    delete L;
  }
  s->lanes.clear();
}

extern "C" int jgpm_serve_start(jgpm_ctx_t ctx, const jgpm_serve_config* cfg, jgpm_serve_t* out) {
  if (!ctx || !out) return JGPM_E_INVALID;
  auto* s = new (std::nothrow) jgpm_serve_s();
  if (!s) return JGPM_E_NOMEM;
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = ctx->hbm / 2;
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 256;
  if (s->cfg.heavy_lanes <= 0) s->cfg.heavy_lanes = 4;
  s->cfg.heavy_lanes = std::min(s->cfg.heavy_lanes, 16);
  for (int i = 0; i < s->cfg.heavy_lanes; ++i) {
    auto* L = new (std::nothrow) Lane();
    int st = L ? jgpm_init(ctx->dev, &L->ctx) : JGPM_E_NOMEM;
    if (st) {
      delete L;
      stop_lanes(s);
      delete s;
      return st;
    }
    L->ctx->tune = ctx->tune;
    s->lanes.push_back(L);
    L->th = std::thread(lane_loop, s, L);
  }
  s->th = std::thread(serve_loop, s);
  *out = s;
  return JGPM_OK;
}

extern "C" int jgpm_serve_submit(jgpm_serve_t s, jgpm_graph_t g, int pattern, int64_t rb, int64_t re,
                                 uint64_t* ticket) {
  if (!s || !g || !ticket || g->ctx != s->ctx) return JGPM_E_INVALID;
  if (pattern < 0 || pattern >= JGPM_NUM_PATTERNS) return JGPM_E_PATTERN;
  if (!jgpm::batch_supported(pattern)) return JGPM_E_PATTERN;  // served ids: TC, P1..P16
  const int64_t nv = g->dag.nv;
  if (re < 0) re = nv;
  if (rb < 0 || rb > re || re > nv) return JGPM_E_INVALID;
  if (!is_dag_pattern(pattern) && !g->has_sym) return JGPM_E_GRAPHKIND;
  // roots wider than 2048 only run in the single-query kernels
  bool wide = false;
  if (is_dag_pattern(pattern) && !g->huge_h.empty()) {
    auto lo = std::lower_bound(g->huge_h.begin(), g->huge_h.end(), (int32_t)rb);
    wide = lo != g->huge_h.end() && *lo < re;
  }
  // grid of the batched kernel (its occupancy is known after the first launch)
  const int occ = s->ctx->batch_occ > 0 ? s->ctx->batch_occ : 3;
  const int want = s->ctx->tune.batch_blocks_per_cu > 0 ? s->ctx->tune.batch_blocks_per_cu : occ;
  const int grid = std::max(1, std::min(want, occ)) * s->ctx->num_cu;
  Req r;
  r.g = g;
  r.pattern = pattern;
  r.rb = rb;
  r.re = re;
  r.cur = rb;
  r.footprint = footprint(g, pattern, s->ctx, grid);
  {
    const std::vector<int64_t>& rp = is_dag_pattern(pattern) ? g->dag_rp_h : g->sym_rp_h;
    r.heavy = wide || (double)(rp[re] - rp[rb]) * pattern_weight(pattern) > (double)s->slice_edges;
  }
  r.t_submit = clk::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return JGPM_E_STATE;
  if (s->fatal) return s->fatal;
  if (r.footprint > s->cfg.hbm_budget_bytes) return JGPM_E_BUDGET;
  r.ticket = s->next_ticket++;
  *ticket = r.ticket;
  s->live.insert(r.ticket);
  if (rb == re) {  // empty range: done at once
    s->done[r.ticket] = Done{0, 0.0, JGPM_OK};
    s->cv_done.notify_all();
    return JGPM_OK;
  }
  s->pending.push_back(r);
  s->cv_work.notify_one();
  return JGPM_OK;
}

// A result is handed out once: retrieving it (poll or wait) erases its
// entry, so a long-running server keeps only unretrieved results.  A ticket
// that is unknown or already retrieved is JGPM_E_INVALID.
static int take_result(jgpm_serve_s* s, uint64_t ticket, uint64_t* count, double* latency_us) {
  auto it = s->done.find(ticket);
  const Done d = it->second;
  s->done.erase(it);
  s->live.erase(ticket);
  if (d.status) return d.status;
  *count = d.count;
  if (latency_us) *latency_us = d.latency_us;
  return JGPM_OK;
}

extern "C" int jgpm_serve_poll(jgpm_serve_t s, uint64_t ticket, uint64_t* count, double* latency_us) {
  if (!s || !count) return JGPM_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->live.count(ticket)) return JGPM_E_INVALID;
  if (!s->done.count(ticket)) return JGPM_E_NOTREADY;
  return take_result(s, ticket, count, latency_us);
}

extern "C" int jgpm_serve_wait(jgpm_serve_t s, uint64_t ticket, uint64_t* count, double* latency_us) {
  if (!s || !count) return JGPM_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->live.count(ticket)) return JGPM_E_INVALID;
  s->cv_done.wait(lk, [&] { return s->done.count(ticket) > 0; });
  return take_result(s, ticket, count, latency_us);
}

extern "C" int jgpm_serve_stop(jgpm_serve_t s) {
  if (!s) return JGPM_E_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  stop_lanes(s);
  delete s;
  return JGPM_OK;
}
