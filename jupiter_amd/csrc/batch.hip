// batch.hip — many (graph, pattern, root range) queries in one persistent
// launch.  No reference counterpart: the reference runs one pattern on one
// graph per process (src/main.cc, src/jupiter_gpm.cu:356-908).
//
// Work is cut into chunks: a chunk is a run of roots (clique patterns, DAG)
// or of directed edges (P1..P12, symmetric CSR) of one query.  Every wave of
// the persistent grid takes the next chunk from one device-wide atomic queue
// (chunks of expensive queries first), processes it with the same device
// routines as the single-query kernels, and adds the chunk's count to that
// query's uint64 slot.  Clique roots with out-degree <= 128 run inside the
// wave (one- or two-word bitmap rows in LDS); the few wider roots of a query
// run afterwards in the block-per-root tier kernels, into the same slot.
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
};

constexpr int BWPB = 4;  // independent waves per block

template <int K>
__device__ __forceinline__ void batch_root(const int64_t* rp, const int32_t* ci, const int32_t* ord, int64_t b0,
                                           int d, int ratio, unsigned char* ar, int32_t* wtot, uint64_t& acc,
                                           const uint64_t* lut) {
  if (d <= 64) {
    using LY = Lay<1, 1, K, false, true>;
    process_root<1, 1, K, false, true>(rp, ci, ord, b0, d, ratio, ar, reinterpret_cast<uint64_t*>(ar + LY::R_OFF),
                                       wtot, acc, lut);
  } else {
    using LY = Lay<1, 2, K, false, true>;
    process_root<1, 2, K, false, true>(rp, ci, ord, b0, d, ratio, ar, reinterpret_cast<uint64_t*>(ar + LY::R_OFF),
                                       wtot, acc, lut);
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
  __shared__ uint64_t s_lut[256];
  fill_bit_lut(s_lut, (int)threadIdx.x, 64 * BWPB);
  __syncthreads();
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
    const int64_t u1 = min(Q.ntasks, u0 + Q.chunk);
    uint64_t acc = 0;
    for (int64_t u = u0; u < u1; ++u) {
      if (Q.kind < 100) {
        const int64_t v0 = Q.base + u;
        const int64_t b0 = Q.rp[v0];
        const int d = (int)(Q.rp[v0 + 1] - b0);
        if (d < Q.kind - 1 || d > 128) continue;
        switch (Q.kind) {
          case 3: batch_root<3>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc, s_lut); break;
          case 4: batch_root<4>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc, s_lut); break;
          case 5: batch_root<5>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc, s_lut); break;
          case 6: batch_root<6>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc, s_lut); break;
          case 7: batch_root<7>(Q.rp, Q.ci, Q.ord, b0, d, a.ratio, ar, s_wtot[wv], acc, s_lut); break;
        }
      } else {
        const int64_t e = Q.base + u;
        const int32_t v0 = Q.src[e], v1 = Q.ci[e];
        int64_t v = 0;
        switch (Q.kind - 100) {
          case 1: v = batch_edge<1>(Q, w, v0, v1, lds_lists, glist, a.scr_stride); break;
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
      hipMalloc(&c->d_bcs, sizeof(int64_t) * (cap + 1)) != hipSuccess ||
      hipMalloc(&c->d_bcnt, sizeof(uint64_t) * cap) != hipSuccess ||
      hipMalloc(&c->d_bq32, sizeof(unsigned) * jk::NUM_TIERS * cap) != hipSuccess)
    return JGPM_E_NOMEM;
  c->bcap = cap;
  return JGPM_OK;
}

__global__ void k_zero_u64(uint64_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_zero_u32(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Enqueue one persistent launch for queries q[0..n) (all kinds supported
// by the kernel), plus the wide-root tier launches.  Counts land in
// c->d_bcnt[i].  Host vectors must stay alive until the stream is synced.
int enqueue_batch(jgpm_ctx_s* c, const jgpm_query* q, int n, std::vector<jk::BQ>& hq, std::vector<int64_t>& cs,
                  hipStream_t st) {
  int rc = ensure_batch_buffers(c, n);
  if (rc) return rc;
  hq.assign(n, jk::BQ{});
  int32_t max_sym_deg = 1;
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
      b.chunk = 16;
    } else {
      const int p = kind - 100;
      rc = ensure_sym_aux(g, pattern_uses_tri(p), p == 3, st);
      if (rc) return rc;
      b.rp = g->sym.rp;
      b.ci = g->sym.ci;
      b.src = g->d_src;
      b.tri = g->d_tri;
      b.nsum = g->d_nsum;
      b.base = g->sym_rp_h[rb];
      b.ntasks = g->sym_rp_h[re] - b.base;
      b.chunk = (p == 2 || p == 3) ? 64 : 8;
      max_sym_deg = std::max(max_sym_deg, g->sym_mmin);
    }
  }
  cs.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) cs[i + 1] = cs[i] + (hq[i].ntasks + hq[i].chunk - 1) / hq[i].chunk;
  if (c->batch_blocks_per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, jk::batch_kernel, 64 * jk::BWPB, 0) != hipSuccess || nb < 1)
      nb = 2;
    c->batch_blocks_per_cu = nb;
  }
  const int grid = std::min(std::max(1, c->tune.blocks_per_cu), c->batch_blocks_per_cu) * c->num_cu;
  const size_t need = (size_t)grid * jk::BWPB * jk::symd::NLIST * (size_t)max_sym_deg;
  if (c->sscr_ints < need) {
    if (hipStreamSynchronize(st) != hipSuccess) return JGPM_E_HIP;
    if (c->d_sscr) (void)hipFree(c->d_sscr);
    c->d_sscr = nullptr;
    c->sscr_ints = 0;
    if (hipMalloc(&c->d_sscr, need * sizeof(int32_t)) != hipSuccess) return JGPM_E_NOMEM;
    c->sscr_ints = need;
  }
  if (hipMemcpyAsync(c->d_bq, hq.data(), sizeof(jk::BQ) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(c->d_bcs, cs.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess)
    return JGPM_E_HIP;
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, st, c->d_bcnt, n);
  hipLaunchKernelGGL(k_zero_u32, dim3((jk::NUM_TIERS * n + 255) / 256), dim3(256), 0, st, c->d_bq32,
                     jk::NUM_TIERS * n);
  jk::BatchArgs a{};
  a.q = (const jk::BQ*)c->d_bq;
  a.cstart = c->d_bcs;
  a.nq = n;
  a.ratio = std::max(1, c->tune.gallop_ratio);
  a.counts = c->d_bcnt;
  a.qctr = c->d_q64 + (CTX_SLOTS - 2);
  a.scr = c->d_sscr;
  a.scr_stride = max_sym_deg;
  prof_begin(c, st, "batch_persistent");
  hipLaunchKernelGGL(jk::batch_kernel, dim3(grid), dim3(64 * jk::BWPB), 0, st, a);
  prof_end(c, st);
  if (hipGetLastError() != hipSuccess) return JGPM_E_HIP;
  // wide clique roots (out-degree > 128) of each query, into its slot
  for (int i = 0; i < n; ++i) {
    if (hq[i].kind >= 100) continue;
    jgpm_graph_s* g = q[i].graph;
    int64_t rb = q[i].root_begin, re = q[i].root_end < 0 ? g->dag.nv : q[i].root_end;
    rc = enqueue_clique(g, hq[i].kind, false, rb, re, 0, st, false, jk::TIER_LDS4, c->d_bcnt + i,
                        c->d_bq32 + (size_t)jk::NUM_TIERS * i);
    if (rc) return rc;
  }
  return JGPM_OK;
}

bool batch_supported(int pattern) { return batch_kind(pattern) != 0; }

int enqueue_batch_q(jgpm_ctx_s* c, const jgpm_query* q, int n, hipStream_t st) {
  // the staging lives on the context until the next batched launch
  return enqueue_batch(c, q, n, c->bq_host, c->bcs_host, st);
}

}  // namespace jgpm

extern "C" int jgpm_count_batch(jgpm_ctx_t ctx, const jgpm_query* q, int n, uint64_t* counts, double* kernel_ms) {
  using namespace jgpm;
  if (!ctx || n < 0 || (n > 0 && (!q || !counts))) return JGPM_E_INVALID;
  for (int i = 0; i < n; ++i) {
    if (!q[i].graph || q[i].graph->ctx != ctx) return JGPM_E_INVALID;
    if (q[i].pattern < 0 || q[i].pattern >= JGPM_NUM_PATTERNS) return JGPM_E_PATTERN;
    const int64_t nv = q[i].graph->dag.nv;
    const int64_t re = q[i].root_end < 0 ? nv : q[i].root_end;
    if (q[i].root_begin < 0 || q[i].root_begin > re || re > nv) return JGPM_E_INVALID;
    const int kind = batch_kind(q[i].pattern);
    if (kind >= 100 && !q[i].graph->has_sym) return JGPM_E_GRAPHKIND;
    if (kind && kind < 100 && !q[i].graph->huge_h.empty()) {
      auto& hv = q[i].graph->huge_h;
      auto lo = std::lower_bound(hv.begin(), hv.end(), (int32_t)q[i].root_begin);
      if (lo != hv.end() && *lo < re) return JGPM_E_BUDGET;
    }
  }
  if (n == 0) {
    if (kernel_ms) *kernel_ms = 0;
    return JGPM_OK;
  }
  // kernel-supported queries, most expensive first
  std::vector<int> idx, rest;
  for (int i = 0; i < n; ++i) (batch_kind(q[i].pattern) ? idx : rest).push_back(i);
  auto cost = [&](int i) {
    const jgpm_graph_s* g = q[i].graph;
    const int kind = batch_kind(q[i].pattern);
    return (double)(kind < 100 ? g->dag.nnz : g->sym.nnz) * (kind < 100 ? kind : 8);
  };
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return cost(x) > cost(y); });
  std::vector<jgpm_query> qs;
  for (int i : idx) qs.push_back(q[i]);
  double total = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->dev) != hipSuccess) return JGPM_E_HIP;
    if (hipEventRecord(ctx->e0, ctx->st) != hipSuccess) return JGPM_E_HIP;
    int st = ctx_reset_counters(ctx, ctx->st);
    if (st) return st;
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
    int st = jgpm_count(q[i].graph, q[i].pattern, q[i].root_begin, q[i].root_end, &counts[i], &m);
    if (st) return st;
    total += m;
  }
  if (kernel_ms) *kernel_ms = total;
  return JGPM_OK;
}
