// batch.hip — many (graph, pattern, root range) queries in one persistent
// launch.  No reference counterpart: the reference runs one pattern on one
// graph per process (src/main.cc, src/jupiter_gpm.cu:356-908).
//
// Work is cut into chunks: a chunk is a run of roots (clique patterns, DAG)
// or of directed edges (P2..P12, symmetric CSR) of one query.  Every wave of
// the persistent grid takes the next chunk from one device-wide atomic queue
// (chunks of expensive queries first), processes it with the same device
// routines as the single-query kernels, and adds the chunk's count to that
// query's uint64 slot.  Clique roots with out-degree up to the hub threshold
// (64 or 128, a tuning knob) run inside the wave (one- or two-word bitmap rows
// in LDS); the few wider roots of a query run afterwards in the block-per-root
// tier kernels, into the same slot, and 4-cycle (P1) queries run afterwards as
// wedge launches (one wave per root), also into their slots.  All launches of a batch can be captured
// once in a HIP graph (prepared batches) and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "clique_dev.hpp"
#include "jgpm_device.hpp"
#include "sym_dev.hpp"

namespace jk {

struct BatchArgs {
  const BQ* q;
  const int64_t* cstart;  // chunk prefix, nq + 1 entries
  int32_t nq;
  int32_t ratio;
  uint64_t* counts;
  unsigned long long* qctr;
  int32_t* scr;           // per-wave list scratch of the symmetric patterns
  int64_t scr_stride;
  int32_t hub;            // clique roots above this out-degree are left to the hub launches (<= 128)
};

constexpr int BWPB = 4;  // independent waves per block

template <int K>
__device__ __forceinline__ void batch_root(const int64_t* rp, const int32_t* ci, const int32_t* ord, int64_t b0,
                                           int d, int ratio, unsigned char* ar, int32_t* wtot, uint64_t& acc) {
  if (d <= 64) {
    using LY = Lay<1, 1, K, false, true>;
    process_root<1, 1, K, false, true>(rp, ci, ord, b0, d, ratio, ar, reinterpret_cast<uint64_t*>(ar + LY::R_OFF),
                                       wtot, acc);
  } else {
    using LY = Lay<1, 2, K, false, true>;
    process_root<1, 2, K, false, true>(rp, ci, ord, b0, d, ratio, ar, reinterpret_cast<uint64_t*>(ar + LY::R_OFF),
                                       wtot, acc);
  }
}

template <int P>
__device__ __forceinline__ int64_t batch_edge(const BQ& Q, symd::Wave& w, int32_t v0, int32_t v1, int32_t* lds,
                                              int32_t* glist, int64_t stride) {
  symd::Ctx g{Q.rp, Q.ci, Q.tri, Q.nsum};
  const int d0 = (int)(Q.rp[v0 + 1] - Q.rp[v0]);
  const int dmin = min(d0, (int)(Q.rp[v1 + 1] - Q.rp[v1]));
  const bool small = dmin <= symd::CAP && ((P != 8 && P != 10) || d0 <= symd::CAP);
  int32_t* L[symd::NLIST];
#pragma unroll
  for (int k = 0; k < symd::NLIST; ++k) L[k] = small ? lds + k * symd::CAP : glist + (int64_t)k * stride;
  return symd::edge_value<P>(g, w, v0, v1, L);
}

template <int A, int B>
struct MaxI {
  static constexpr int v = A > B ? A : B;
};

__global__ void __launch_bounds__(64 * BWPB) batch_kernel(BatchArgs a) {
  constexpr int CL = MaxI<MaxI<Lay<1, 2, 7, false, true>::SLOT, Lay<1, 2, 5, false, true>::SLOT>::v,
                          MaxI<Lay<1, 2, 4, false, true>::SLOT, Lay<1, 1, 7, false, true>::SLOT>::v>::v;
  constexpr int SY = (symd::NLIST * symd::CAP + symd::STAGE) * 4;
  constexpr int ARENA = ((CL > SY ? CL : SY) + 15) & ~15;
  __shared__ __attribute__((aligned(16))) unsigned char s_arena[BWPB][ARENA];
  __shared__ int32_t s_wtot[BWPB][2];
  const int wv = threadIdx.x >> 6;
  const int ln = lane_id();
  unsigned char* ar = s_arena[wv];
  int32_t* lds_lists = reinterpret_cast<int32_t*>(ar);
  symd::Wave w{lds_lists + symd::NLIST * symd::CAP};
  int32_t* glist = a.scr + ((int64_t)blockIdx.x * BWPB + wv) * symd::NLIST * a.scr_stride;
  const int64_t nchunks = a.cstart[a.nq];
  for (;;) {
    int64_t t = 0;
    if (ln == 0) t = (int64_t)atomicAdd(a.qctr, 1ull);
    const int64_t k = __shfl(t, 0);
    if (k >= nchunks) break;
    // query of chunk k: last q with cstart[q] <= k
    int lo = 0, hi = a.nq;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (a.cstart[mid] <= k) lo = mid; else hi = mid;
    }
    const BQ Q = a.q[lo];
    const int64_t u0 = (k - a.cstart[lo]) * Q.chunk;
    int64_t u1 = min(Q.ntasks, u0 + Q.chunk);
    uint64_t acc = 0;
    if (Q.kind == 102 && Q.tri) {
      // diamonds from the per-edge triangle array: lanes over the edges
      for (int64_t u = u0 + ln; u < u1; u += 64) {
        const int64_t e = Q.base + u;
        if (Q.ci[e] < Q.src[e]) acc += (uint64_t)symd::C2(Q.tri[e]);
      }
      u1 = u0;  // no per-edge pass
    }
    for (int64_t u = u0; u < u1; ++u) {
      if (Q.kind < 100) {
        const int64_t v0 = Q.base + u;
        const int64_t b0 = Q.rp[v0];
        const int d = (int)(Q.rp[v0 + 1] - b0);
        if (d < Q.kind - 1 || d > a.hub) continue;
        switch (Q.kind) {
          case 3: batch_root<3>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc); break;
          case 4: batch_root<4>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc); break;
          case 5: batch_root<5>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc); break;
          case 6: batch_root<6>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc); break;
          case 7: batch_root<7>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc); break;
        }
      } else {
        const int64_t e = Q.base + u;
        const int32_t v0 = Q.src[e], v1 = Q.ci[e];
        int64_t v = 0;
        switch (Q.kind - 100) {
          case 2: v = batch_edge<2>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 3: v = batch_edge<3>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 4: v = batch_edge<4>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 5: v = batch_edge<5>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 6: v = batch_edge<6>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 7: v = batch_edge<7>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 8: v = batch_edge<8>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 9: v = batch_edge<9>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 10: v = batch_edge<10>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 11: v = batch_edge<11>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
          case 12: v = batch_edge<12>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
        }
        if (ln == 0) acc += (uint64_t)v;
      }
    }
    acc = wave_sum_u64(acc);
    if (ln == 0 && acc) atomicAdd((unsigned long long*)&a.counts[lo], (unsigned long long)acc);
  }
}

}  // namespace jk

namespace jgpm {

// Kinds the persistent kernel runs; other patterns go through the
// single-query path after it.
static int batch_kind(int pattern) {
  int k = clique_k(pattern);
  if (k && pattern != JGPM_I_K4) return k;
  if (pattern >= JGPM_P1 && pattern <= JGPM_P12) return 100 + pattern;
  return 0;
}

bool batch_supported(int pattern) { return batch_kind(pattern) != 0; }

// Device buffers one batched launch reads and writes.
struct BatchBufs {
  jk::BQ* bq = nullptr;
  int64_t* bcs = nullptr;
  uint64_t* bcnt = nullptr;
  unsigned* bq32 = nullptr;             // queue counters of the hub launches, NUM_TIERS per query
  unsigned long long* qctr = nullptr;   // chunk queue of the persistent kernel
  int32_t* sscr = nullptr;              // list scratch, grid * BWPB * NLIST * stride ints
  int64_t sscr_stride = 1;
  uint64_t* gscr = nullptr;             // global-row tier scratch (nullptr: the context's)
  size_t gscr_words = 0;
  int32_t* p1cnt = nullptr;             // zeroed wedge counters of the 4-cycle launches
  size_t p1cnt_ints = 0;
};

__global__ void k_zero_u64(uint64_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_zero_u32(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

static int batch_grid(jgpm_ctx_s* c) {
  if (c->batch_occ == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, jk::batch_kernel, 64 * jk::BWPB, 0) != hipSuccess || nb < 1)
      nb = 2;
    c->batch_occ = nb;
  }
  const int want = c->tune.batch_blocks_per_cu > 0 ? c->tune.batch_blocks_per_cu : c->batch_occ;
  return std::min(want, c->batch_occ) * c->num_cu;
}

// Host side of a batch: the query table (work items sized by the tuning),
// its chunk prefix and the list-scratch stride.  May build a graph's lazy
// auxiliary arrays on st (kernels and allocations), so it is never captured.
static int plan_batch(jgpm_ctx_s* c, const jgpm_query* q, int n, std::vector<jk::BQ>& hq, std::vector<int64_t>& cs,
                      int32_t& max_sym, size_t& p1_ints, hipStream_t st) {
  // cs holds two prefixes of n + 1 entries: the persistent kernel's chunks,
  // then the roots of the 4-cycle queries (one batched wedge launch)
  const int rchunk = std::max(1, c->tune.batch_root_chunk);
  const int echunk = std::max(1, c->tune.batch_edge_chunk);
  hq.assign(n, jk::BQ{});
  max_sym = 1;
  p1_ints = 0;
  for (int i = 0; i < n; ++i) {
    jgpm_graph_s* g = q[i].graph;
    const int kind = batch_kind(q[i].pattern);
    int64_t rb = q[i].root_begin, re = q[i].root_end < 0 ? g->dag.nv : q[i].root_end;
    jk::BQ& b = hq[i];
    b.kind = kind;
    if (kind < 100) {
      b.rp = g->dag.rp;
      b.ci = g->dag.ci;
      b.ord = g->d_ord;
      b.base = rb;
      b.ntasks = re - rb;
      b.chunk = rchunk;
    } else if (kind == 101) {
      // 4-cycles run as wedge launches after the persistent kernel (no chunks here)
      p1_ints = std::max(p1_ints, p1_counter_ints(c, (int32_t)g->sym.nv));
      const bool whole = rb == 0 && re == g->sym.nv;
      if (whole) {
        int rc = ensure_rank_csr(g, st);
        if (rc) return rc;
      }
      // a whole-graph count runs on the degree-ordered relabelling (same
      // total, hubs ranked last), as the single-query launch does
      b.rp = whole && g->rsym.ci ? g->rsym.rp : g->sym.rp;
      b.ci = whole && g->rsym.ci ? g->rsym.ci : g->sym.ci;
      b.base = rb;
      b.ntasks = 0;
      b.chunk = 1;
    } else {
      const int p = kind - 100;
      // diamonds read the per-edge triangle array, as the single-query path does
      int rc = ensure_sym_aux(g, p == 2 || pattern_uses_tri(p), p == 3, st);
      if (rc) return rc;
      b.rp = g->sym.rp;
      b.ci = g->sym.ci;
      b.src = g->d_src;
      b.tri = g->d_tri;
      b.nsum = g->d_nsum;
      b.base = g->sym_rp_h[rb];
      b.ntasks = g->sym_rp_h[re] - b.base;
      b.chunk = p == 2 ? 64 * echunk : p == 3 ? 8 * echunk : echunk;
      max_sym = std::max(max_sym, g->sym_mmin);
    }
  }
  cs.assign(2 * (size_t)(n + 1), 0);
  for (int i = 0; i < n; ++i) cs[i + 1] = cs[i] + (hq[i].ntasks + hq[i].chunk - 1) / hq[i].chunk;
  int64_t* p1cs = cs.data() + n + 1;
  for (int i = 0; i < n; ++i) {
    int64_t r = 0;
    if (hq[i].kind == 101) {
      const jgpm_graph_s* g = q[i].graph;
      r = (q[i].root_end < 0 ? g->sym.nv : q[i].root_end) - q[i].root_begin;
    }
    p1cs[i + 1] = p1cs[i] + r;
  }
  return JGPM_OK;
}

// Every launch of one batch, and nothing else (no allocation, host copy or
// memset), so that the sequence can be captured: counter fills, the
// persistent kernel, then the hub launches of each clique query into its slot.
static int launch_batch(jgpm_ctx_s* c, const jgpm_query* q, int n, const std::vector<jk::BQ>& hq, const BatchBufs& b,
                        int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, st, b.bcnt, n);
  hipLaunchKernelGGL(k_zero_u32, dim3((jk::NUM_TIERS * n + 255) / 256), dim3(256), 0, st, b.bq32, jk::NUM_TIERS * n);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, reinterpret_cast<uint64_t*>(b.qctr), 1);
  const int hub = c->tune.batch_hub_deg <= 64 ? 64 : 128;
  jk::BatchArgs a{};
  a.q = b.bq;
  a.cstart = b.bcs;
  a.nq = n;
  a.ratio = std::max(1, c->tune.gallop_ratio);
  a.counts = b.bcnt;
  a.qctr = b.qctr;
  a.scr = b.sscr;
  a.scr_stride = b.sscr_stride;
  a.hub = hub;
  prof_begin(c, st, "batch_persistent");
  hipLaunchKernelGGL(jk::batch_kernel, dim3(grid), dim3(64 * jk::BWPB), 0, st, a);
  prof_end(c, st);
  if (hipGetLastError() != hipSuccess) return JGPM_E_HIP;
  const int first_tier = hub <= 64 ? jk::TIER_LDS2 : jk::TIER_LDS4;
  // every 4-cycle query in one wedge launch over their flattened roots
  {
    int first = -1;
    int64_t total = 0;
    int32_t max_nv = 1;
    for (int i = 0; i < n; ++i)
      if (hq[i].kind == 101) {
        if (first < 0) first = i;
        const jgpm_graph_s* g = q[i].graph;
        total += (q[i].root_end < 0 ? g->sym.nv : q[i].root_end) - q[i].root_begin;
        max_nv = std::max(max_nv, (int32_t)g->sym.nv);
      }
    if (total > 0) {
      const int64_t waves = std::min<int64_t>((int64_t)c->num_cu * 32, (int64_t)(b.p1cnt_ints / (size_t)max_nv));
      const int grid = (int)std::min<int64_t>(waves / jk::P1_WPB, (total + jk::P1_WPB - 1) / jk::P1_WPB);
      if (!b.p1cnt || grid < 1) return JGPM_E_STATE;  // counters not sized for these graphs
      jk::P1BatchArgs pa{};
      pa.q = b.bq;
      pa.p1cs = b.bcs + (n + 1);
      pa.nq = n;
      pa.total = total;
      pa.qctr = b.bq32 + (size_t)jk::NUM_TIERS * first;  // zeroed above, unused by the wedge query otherwise
      pa.cnt = b.p1cnt;
      pa.cnt_stride = max_nv;
      pa.counts = b.bcnt;
      prof_begin(c, st, "batch_p1_wedge");
      hipError_t e = jk::launch_p1_batch(pa, grid, st);
      prof_end(c, st);
      if (e != hipSuccess) return JGPM_E_HIP;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (hq[i].kind >= 100) continue;
    jgpm_graph_s* g = q[i].graph;
    int64_t rb = q[i].root_begin, re = q[i].root_end < 0 ? g->dag.nv : q[i].root_end;
    int rc = enqueue_clique(g, hq[i].kind, false, rb, re, 0, st, false, first_tier, b.bcnt + i,
                            b.bq32 + (size_t)jk::NUM_TIERS * i, b.gscr, b.gscr_words);
    if (rc) return rc;
  }
  return JGPM_OK;
}

static int ensure_batch_buffers(jgpm_ctx_s* c, int n) {
  if (c->bcap >= n) return JGPM_OK;
  if (hipStreamSynchronize(c->st) != hipSuccess) return JGPM_E_HIP;
  if (c->d_bq) (void)hipFree(c->d_bq);
  if (c->d_bcs) (void)hipFree(c->d_bcs);
  if (c->d_bcnt) (void)hipFree(c->d_bcnt);
  if (c->d_bq32) (void)hipFree(c->d_bq32);
  c->d_bq = nullptr;
  c->d_bcs = nullptr;
  c->d_bcnt = nullptr;
  c->d_bq32 = nullptr;
  c->bcap = 0;
  int cap = std::max(64, n);
  if (hipMalloc(&c->d_bq, sizeof(jk::BQ) * cap) != hipSuccess ||
      hipMalloc(&c->d_bcs, sizeof(int64_t) * 2 * (cap + 1)) != hipSuccess ||
      hipMalloc(&c->d_bcnt, sizeof(uint64_t) * cap) != hipSuccess ||
      hipMalloc(&c->d_bq32, sizeof(unsigned) * jk::NUM_TIERS * cap) != hipSuccess)
    return JGPM_E_NOMEM;
  c->bcap = cap;
  return JGPM_OK;
}

// Eager batched launch on the context's buffers for queries q[0..n) (all of
// kinds the kernel runs).  Counts land in c->d_bcnt[i].  hq / cs must stay
// alive until the stream is synchronised.
int enqueue_batch(jgpm_ctx_s* c, const jgpm_query* q, int n, std::vector<jk::BQ>& hq, std::vector<int64_t>& cs,
                  hipStream_t st) {
  int rc = ensure_batch_buffers(c, n);
  if (rc) return rc;
  int32_t max_sym = 1;
  size_t p1_ints = 0;
  rc = plan_batch(c, q, n, hq, cs, max_sym, p1_ints, st);
  if (rc) return rc;
  if (p1_ints && (rc = ensure_p1_counters(c, p1_ints, st))) return rc;
  const int grid = batch_grid(c);
  const size_t need = (size_t)grid * jk::BWPB * jk::symd::NLIST * (size_t)max_sym;
  if (c->sscr_ints < need) {
    if (hipStreamSynchronize(st) != hipSuccess) return JGPM_E_HIP;
    if (c->d_sscr) (void)hipFree(c->d_sscr);
    c->d_sscr = nullptr;
    c->sscr_ints = 0;
    if (hipMalloc(&c->d_sscr, need * sizeof(int32_t)) != hipSuccess) return JGPM_E_NOMEM;
    c->sscr_ints = need;
  }
  if (hipMemcpyAsync(c->d_bq, hq.data(), sizeof(jk::BQ) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(c->d_bcs, cs.data(), sizeof(int64_t) * 2 * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess)
    return JGPM_E_HIP;
  BatchBufs b;
  b.bq = static_cast<jk::BQ*>(c->d_bq);
  b.bcs = c->d_bcs;
  b.bcnt = c->d_bcnt;
  b.p1cnt = c->d_p1cnt;
  b.p1cnt_ints = c->p1cnt_ints;
  b.bq32 = c->d_bq32;
  b.qctr = c->d_q64 + (CTX_SLOTS - 2);
  b.sscr = c->d_sscr;
  b.sscr_stride = max_sym;
  return launch_batch(c, q, n, hq, b, grid, st);
}

int enqueue_batch_q(jgpm_ctx_s* c, const jgpm_query* q, int n, hipStream_t st) {
  // the staging lives on the context until the next batched launch
  return enqueue_batch(c, q, n, c->bq_host, c->bcs_host, st);
}

// Checks shared by the batched entry points; `kernel_only` rejects patterns
// the persistent kernel does not run.
static int check_queries(jgpm_ctx_s* ctx, const jgpm_query* q, int n, bool kernel_only) {
  for (int i = 0; i < n; ++i) {
    if (!q[i].graph || q[i].graph->ctx != ctx) return JGPM_E_INVALID;
    if (q[i].pattern < 0 || q[i].pattern >= JGPM_NUM_PATTERNS) return JGPM_E_PATTERN;
    const int64_t nv = q[i].graph->dag.nv;
    const int64_t re = q[i].root_end < 0 ? nv : q[i].root_end;
    if (q[i].root_begin < 0 || q[i].root_begin > re || re > nv) return JGPM_E_INVALID;
    const int kind = batch_kind(q[i].pattern);
    if (kernel_only && !kind) return JGPM_E_PATTERN;
    if (kind >= 100 && !q[i].graph->has_sym) return JGPM_E_GRAPHKIND;
    if (kind && kind < 100 && !q[i].graph->huge_h.empty()) {
      auto& hv = q[i].graph->huge_h;
      auto lo = std::lower_bound(hv.begin(), hv.end(), (int32_t)q[i].root_begin);
      if (lo != hv.end() && *lo < re) return JGPM_E_BUDGET;
    }
  }
  return JGPM_OK;
}

// Kernel-run queries in launch order: most expensive first.
static std::vector<int> cost_order(const jgpm_query* q, const std::vector<int>& idx_in) {
  std::vector<int> idx = idx_in;
  auto cost = [&](int i) {
    const jgpm_graph_s* g = q[i].graph;
    const int kind = batch_kind(q[i].pattern);
    return (double)(kind < 100 ? g->dag.nnz : g->sym.nnz) * (kind < 100 ? kind : 8);
  };
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return cost(x) > cost(y); });
  return idx;
}

}  // namespace jgpm

extern "C" int jgpm_count_batch(jgpm_ctx_t ctx, const jgpm_query* q, int n, uint64_t* counts, double* kernel_ms) {
  using namespace jgpm;
  if (!ctx || n < 0 || (n > 0 && (!q || !counts))) return JGPM_E_INVALID;
  int st = check_queries(ctx, q, n, false);
  if (st) return st;
  if (n == 0) {
    if (kernel_ms) *kernel_ms = 0;
    return JGPM_OK;
  }
  std::vector<int> all, rest;
  for (int i = 0; i < n; ++i) (batch_kind(q[i].pattern) ? all : rest).push_back(i);
  const std::vector<int> idx = cost_order(q, all);
  std::vector<jgpm_query> qs;
  for (int i : idx) qs.push_back(q[i]);
  double total = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->dev) != hipSuccess) return JGPM_E_HIP;
    if (hipEventRecord(ctx->e0, ctx->st) != hipSuccess) return JGPM_E_HIP;
    ctx->nprof = 0;
    std::vector<jk::BQ> hq;
    std::vector<int64_t> cs;
    if (!qs.empty()) {
      st = enqueue_batch(ctx, qs.data(), (int)qs.size(), hq, cs, ctx->st);
      if (st) {
        (void)hipStreamSynchronize(ctx->st);
        return st;
      }
    }
    std::vector<uint64_t> h(qs.size());
    if (!qs.empty() && hipMemcpyAsync(h.data(), ctx->d_bcnt, sizeof(uint64_t) * qs.size(), hipMemcpyDeviceToHost,
                                      ctx->st) != hipSuccess)
      return JGPM_E_HIP;
    if (hipEventRecord(ctx->e1, ctx->st) != hipSuccess || hipStreamSynchronize(ctx->st) != hipSuccess)
      return JGPM_E_HIP;
    for (size_t k = 0; k < idx.size(); ++k) counts[idx[k]] = h[k];
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->e0, ctx->e1);
    total = ms;
  }
  // composite patterns (induced motifs, vertex sums) run one by one
  for (int i : rest) {
    double m = 0;
    st = jgpm_count(q[i].graph, q[i].pattern, q[i].root_begin, q[i].root_end, &counts[i], &m);
    if (st) return st;
    total += m;
  }
  if (kernel_ms) *kernel_ms = total;
  return JGPM_OK;
}

// ---------------------------------------------------------------------------
// Prepared batches: own buffers, one captured HIP graph, replay per run.

struct jgpm_batch_s {
  jgpm_ctx_s* ctx = nullptr;
  std::vector<jgpm_query> q;   // launch order
  std::vector<int> idx;        // launch slot -> caller index
  std::vector<jk::BQ> hq;
  std::vector<int64_t> cs;
  jgpm::BatchBufs b;
  int grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> h;
};

static void batch_release(jgpm_batch_s* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  void* bufs[] = {B->b.bq, B->b.bcs, B->b.bcnt, B->b.bq32, B->b.qctr, B->b.sscr, B->b.gscr, B->b.p1cnt};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (const jgpm_query& x : B->q) x.graph->nbatch--;
  delete B;
}

extern "C" int jgpm_batch_create(jgpm_ctx_t ctx, const jgpm_query* q, int n, jgpm_batch_t* out) {
  using namespace jgpm;
  if (!ctx || !out || n <= 0 || !q) return JGPM_E_INVALID;
  int st = check_queries(ctx, q, n, true);
  if (st) return st;
  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  auto* B = new (std::nothrow) jgpm_batch_s();
  if (!B) return JGPM_E_NOMEM;
  B->ctx = ctx;
  B->idx = cost_order(q, all);
  for (int i : B->idx) {
    B->q.push_back(q[i]);
    q[i].graph->nbatch++;
  }
  B->h.assign(n, 0);
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(ctx->st);
    batch_release(B);
    return code;
  };
  if (hipSetDevice(ctx->dev) != hipSuccess) return fail(JGPM_E_HIP);
  int32_t max_sym = 1;
  size_t p1_ints = 0;
  st = plan_batch(ctx, B->q.data(), n, B->hq, B->cs, max_sym, p1_ints, ctx->st);
  if (st) return fail(st);
  B->grid = batch_grid(ctx);
  // global-row scratch only when a clique query has roots in that tier
  bool glb = false;
  for (int i = 0; i < n; ++i) {
    if (B->hq[i].kind >= 100) continue;
    const auto& v = B->q[i].graph->tier_h[jk::TIER_GLB32];
    const int64_t rb = B->q[i].root_begin, re = B->q[i].root_end < 0 ? B->q[i].graph->dag.nv : B->q[i].root_end;
    auto lo = std::lower_bound(v.begin(), v.end(), (int32_t)rb);
    glb = glb || (lo != v.end() && *lo < re);
  }
  B->b.sscr_stride = max_sym;
  const size_t sscr = (size_t)B->grid * jk::BWPB * jk::symd::NLIST * (size_t)max_sym;
  if (glb) B->b.gscr_words = jk::tier_scratch_words(jk::TIER_GLB32) * (size_t)glb_tier_grid(ctx);
  if (hipMalloc(&B->b.bq, sizeof(jk::BQ) * n) != hipSuccess ||
      hipMalloc(&B->b.bcs, sizeof(int64_t) * 2 * (n + 1)) != hipSuccess ||
      hipMalloc(&B->b.bcnt, sizeof(uint64_t) * n) != hipSuccess ||
      hipMalloc(&B->b.bq32, sizeof(unsigned) * jk::NUM_TIERS * n) != hipSuccess ||
      hipMalloc(&B->b.qctr, 64) != hipSuccess || hipMalloc(&B->b.sscr, sscr * sizeof(int32_t)) != hipSuccess ||
      (glb && hipMalloc(&B->b.gscr, B->b.gscr_words * sizeof(uint64_t)) != hipSuccess) ||
      (p1_ints && hipMalloc(&B->b.p1cnt, p1_ints * sizeof(int32_t)) != hipSuccess))
    return fail(JGPM_E_NOMEM);
  // the batch's own wedge counters (a later context allocation cannot move
  // them under the captured launches): zeroed once, left zero by every replay
  B->b.p1cnt_ints = p1_ints;
  if (p1_ints && hipMemsetAsync(B->b.p1cnt, 0, p1_ints * sizeof(int32_t), ctx->st) != hipSuccess)
    return fail(JGPM_E_HIP);
  if (hipMemcpyAsync(B->b.bq, B->hq.data(), sizeof(jk::BQ) * n, hipMemcpyHostToDevice, ctx->st) != hipSuccess ||
      hipMemcpyAsync(B->b.bcs, B->cs.data(), sizeof(int64_t) * 2 * (n + 1), hipMemcpyHostToDevice, ctx->st) !=
          hipSuccess ||
      hipStreamSynchronize(ctx->st) != hipSuccess)
    return fail(JGPM_E_HIP);
  // capture every launch of the batch (no allocations, host copies or
  // memsets happen inside launch_batch; profiling events stay out)
  const bool prof = ctx->prof_on;
  ctx->prof_on = false;
  ctx->capturing = true;
  if (hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    ctx->prof_on = prof;
    ctx->capturing = false;
    return fail(JGPM_E_HIP);
  }
  st = launch_batch(ctx, B->q.data(), n, B->hq, B->b, B->grid, ctx->st);
  const hipError_t ce = hipStreamEndCapture(ctx->st, &B->graph);
  ctx->prof_on = prof;
  ctx->capturing = false;
  if (st) return fail(st);
  if (ce != hipSuccess || !B->graph) return fail(JGPM_E_HIP);
  if (hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0) != hipSuccess) return fail(JGPM_E_HIP);
  *out = B;
  return JGPM_OK;
}

extern "C" int jgpm_batch_run(jgpm_batch_t B, uint64_t* counts, double* kernel_ms) {
  if (!B || !counts) return JGPM_E_INVALID;
  jgpm_ctx_s* ctx = B->ctx;
  const int n = (int)B->q.size();
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->dev) != hipSuccess) return JGPM_E_HIP;
  ctx->nprof = 0;
  if (hipEventRecord(ctx->e0, ctx->st) != hipSuccess || hipGraphLaunch(B->exec, ctx->st) != hipSuccess ||
      hipEventRecord(ctx->e1, ctx->st) != hipSuccess ||
      hipMemcpyAsync(B->h.data(), B->b.bcnt, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, ctx->st) != hipSuccess ||
      hipStreamSynchronize(ctx->st) != hipSuccess)
    return JGPM_E_HIP;
  for (int k = 0; k < n; ++k) counts[B->idx[k]] = B->h[k];
  if (kernel_ms) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->e0, ctx->e1);
    *kernel_ms = ms;
  }
  return JGPM_OK;
}

extern "C" int jgpm_batch_free(jgpm_batch_t B) {
  if (!B) return JGPM_E_INVALID;
  {
    std::lock_guard<std::mutex> lk(B->ctx->mu);
    (void)hipSetDevice(B->ctx->dev);
    (void)hipStreamSynchronize(B->ctx->st);
  }
  batch_release(B);
  return JGPM_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: coordinate descent over the batched-launch knobs.

extern "C" int jgpm_autotune_batch(jgpm_ctx_t ctx, const jgpm_query* q, int n, int reps, jgpm_tuning* best,
                                   double* best_ms) {
  using namespace jgpm;
  if (!ctx || n <= 0 || !q) return JGPM_E_INVALID;
  int st = check_queries(ctx, q, n, false);
  if (st) return st;
  reps = std::max(1, std::min(reps, 16));
  const jgpm_tuning start = ctx->tune;
  std::vector<uint64_t> ref(n), got(n);
  bool have_ref = false;
  auto eval = [&](const jgpm_tuning& t, double& ms) -> int {
    ctx->tune = t;
    ms = 1e300;
    for (int r = 0; r < reps; ++r) {
      double m = 0;
      int rc = jgpm_count_batch(ctx, q, n, got.data(), &m);
      if (rc) return rc;
      ms = std::min(ms, m);
      if (!have_ref) {
        ref = got;
        have_ref = true;
      } else if (got != ref) {
        return JGPM_E_STATE;  // every setting must give the same exact counts
      }
    }
    return JGPM_OK;
  };
  double cur_ms = 0, m = 0;
  st = eval(start, m);  // warm-up: builds lazy per-graph arrays
  if (!st) st = eval(start, cur_ms);
  jgpm_tuning cur = start;
  const int occ = std::max(1, ctx->batch_occ);
  const int hub_opt[] = {64, 128};
  const int rchunk_opt[] = {4, 16, 64};
  const int echunk_opt[] = {2, 8, 32};
  const int bpc_opt[] = {occ, std::max(1, occ / 2)};
  for (int knob = 0; knob < 4 && !st; ++knob) {
    const int* opt = knob == 0 ? hub_opt : knob == 1 ? rchunk_opt : knob == 2 ? echunk_opt : bpc_opt;
    const int nopt = knob == 0 || knob == 3 ? 2 : 3;
    for (int k = 0; k < nopt && !st; ++k) {
      jgpm_tuning t = cur;
      int32_t* f = knob == 0 ? &t.batch_hub_deg : knob == 1 ? &t.batch_root_chunk
                 : knob == 2 ? &t.batch_edge_chunk : &t.batch_blocks_per_cu;
      if (*f == opt[k] || (knob == 3 && *f <= 0 && opt[k] == occ)) continue;
      *f = opt[k];
      st = eval(t, m);
      if (!st && m < cur_ms) {
        cur = t;
        cur_ms = m;
      }
    }
  }
  ctx->tune = st ? start : cur;
  if (st) return st;
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return JGPM_OK;
}
