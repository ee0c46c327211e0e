// dispatch.hip — pattern id -> kernels, and the combination of partial
// accumulators into the final count.
//
// The reference chooses kernels by string compare and asserts on an unknown
// name (src/jupiter_gpm.cu:391-433,435-873); here an id selects an enqueue
// routine and an unknown id is JGPM_E_PATTERN.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jgpm_device.hpp"

namespace jgpm {

static int grid_for(jgpm_ctx_s* c, int per_cu) { return std::max(1, per_cu) * c->num_cu; }

static thread_local jgpm_ctx_s* t_exec = nullptr;

jgpm_ctx_s* exec_ctx(const jgpm_graph_s* g) { return t_exec ? t_exec : g->ctx; }
ExecScope::ExecScope(jgpm_ctx_s* c) : prev(t_exec) { t_exec = c; }
ExecScope::~ExecScope() { t_exec = prev; }

// Lazily built per-graph arrays of the symmetric kernels.
int ensure_sym_aux(jgpm_graph_s* g, bool need_tri, bool need_nsum, hipStream_t st) {
  jgpm_ctx_s* c = exec_ctx(g);
  const int64_t ne = g->sym.nnz;
  if (!g->d_src) {
    if (hipMalloc(&g->d_src, sizeof(int32_t) * std::max<int64_t>(ne, 1)) != hipSuccess) return JGPM_E_NOMEM;
    g->bytes += 4 * ne;
    if (jk::launch_aux_src(g->sym.rp, g->sym.nv, g->d_src, grid_for(c, 4), st) != hipSuccess) return JGPM_E_HIP;
  }
  if (need_tri && !g->d_tri) {
    if (hipMalloc(&g->d_tri, sizeof(int32_t) * std::max<int64_t>(ne, 1)) != hipSuccess) return JGPM_E_NOMEM;
    g->bytes += 4 * ne;
    // its own queue counter, zeroed here: the builder runs from single
    // queries, batch plans and the serving scheduler alike, and not every
    // caller resets the context's counters first
    unsigned long long* q = c->d_q64 + (CTX_SLOTS - 1);
    if (zero_u64(reinterpret_cast<uint64_t*>(q), 1, st)) return JGPM_E_HIP;
    prof_begin(c, st, "edge_triangles");
    hipError_t e = jk::launch_aux_tri(g->sym.rp, g->sym.ci, g->d_src, ne, g->d_tri, q, grid_for(c, 4), st);
    prof_end(c, st);
    if (e != hipSuccess) return JGPM_E_HIP;
  }
  if (need_nsum && !g->d_nsum) {
    if (hipMalloc(&g->d_nsum, sizeof(int64_t) * (g->sym.nv + 1)) != hipSuccess) return JGPM_E_NOMEM;
    g->bytes += 8 * (int64_t)g->sym.nv;
    if (jk::launch_aux_nsum(g->sym.rp, g->sym.ci, g->sym.nv, g->d_nsum, grid_for(c, 4), st) != hipSuccess)
      return JGPM_E_HIP;
  }
  return JGPM_OK;
}

bool pattern_uses_tri(int p) { return p == 3 || p == 4 || p == 7 || p == 10 || p == 11; }

// The single-query launches of P5, P7, P9, P11 and P12 read the triangle
// array to send their heavy edges to the block-per-edge kernel.
static bool heavy_split_on(const jgpm_ctx_s* c, int p) {
  return jk::sym_heavy_pattern(p) && c->tune.sym_heavy_min < jk::sym_heavy_nmax();
}

bool query_needs_tri(const jgpm_ctx_s* c, int p) { return p == 2 || pattern_uses_tri(p) || heavy_split_on(c, p); }

// Waves of the 4-cycle wedge kernel on a graph of nv vertices: one dense
// counter array of nv ints each, within 2 GiB, at most 32 waves per CU.
static int64_t p1_waves(jgpm_ctx_s* c, int32_t nv) {
  const int64_t budget = (int64_t)1 << 29;  // ints
  const int64_t cap = (int64_t)c->num_cu * 32;
  return std::max<int64_t>(jk::P1_WPB, std::min<int64_t>(cap, budget / std::max<int32_t>(nv, 1)));
}

size_t p1_counter_ints(jgpm_ctx_s* c, int32_t nv) { return (size_t)p1_waves(c, nv) * (size_t)std::max<int32_t>(nv, 1); }

int ensure_p1_counters(jgpm_ctx_s* c, size_t need, hipStream_t st) {
  if (c->p1cnt_ints >= need) return JGPM_OK;
  if (c->capturing) return JGPM_E_STATE;
  if (c->d_p1cnt) {
    if (hipStreamSynchronize(st) != hipSuccess) return JGPM_E_HIP;
    (void)hipFree(c->d_p1cnt);
    c->d_p1cnt = nullptr;
    c->p1cnt_ints = 0;
  }
  if (hipMalloc(&c->d_p1cnt, need * sizeof(int32_t)) != hipSuccess) return JGPM_E_NOMEM;
  // zero once: every launch leaves the counters it touched at zero again
  if (hipMemsetAsync(c->d_p1cnt, 0, need * sizeof(int32_t), st) != hipSuccess) return JGPM_E_HIP;
  c->p1cnt_ints = need;
  return JGPM_OK;
}

int launch_p1(jgpm_graph_s* g, int64_t rb, int64_t re, uint64_t* out, unsigned* qctr, int32_t* cnt, size_t cnt_ints,
              hipStream_t st) {
  jgpm_ctx_s* c = exec_ctx(g);
  const int32_t nv = (int32_t)g->sym.nv;
  if (re <= rb) return JGPM_OK;
  const int64_t waves = std::min<int64_t>(p1_waves(c, nv), (int64_t)(cnt_ints / std::max<int32_t>(nv, 1)));
  const int grid = (int)std::min<int64_t>(waves / jk::P1_WPB, (re - rb + jk::P1_WPB - 1) / jk::P1_WPB);
  if (!cnt || grid < 1) return JGPM_E_STATE;  // counters not sized for this graph
  // a whole-graph count runs on the degree-ordered relabelling when it is
  // built: the same identity in another vertex order gives the same total,
  // and the hubs, ranked last, are no wedge's middle vertex
  const bool ranked = rb == 0 && re == nv && g->rsym.ci;
  jk::P1Args a{};
  a.rp = ranked ? g->rsym.rp : g->sym.rp;
  a.ci = ranked ? g->rsym.ci : g->sym.ci;
  a.rb = rb;
  a.re = re;
  a.nv = nv;
  // on the relabelling the hubs rank last: deal them first, so the widest
  // roots do not run alone at the end of the launch
  a.desc = ranked ? 1 : 0;
  a.qctr = qctr;
  a.cnt = cnt;
  a.out = out;
  prof_begin(c, st, "sym_p1_wedge");
  hipError_t e = jk::launch_p1_wedge(a, grid, st);
  prof_end(c, st);
  return e == hipSuccess ? JGPM_OK : JGPM_E_HIP;
}

int enqueue_sym(jgpm_graph_s* g, int p, int64_t rb, int64_t re, int slot, hipStream_t st) {
  if (!g->has_sym) return JGPM_E_GRAPHKIND;
  jgpm_ctx_s* c = exec_ctx(g);
  if (p == 1) {
    int rc = ensure_p1_counters(c, p1_counter_ints(c, (int32_t)g->sym.nv), st);
    if (rc) return rc;
    if (rb == 0 && re == g->sym.nv && (rc = ensure_rank_csr(g, st))) return rc;
    return launch_p1(g, rb, re, c->d_out + slot, c->d_q + slot, c->d_p1cnt, c->p1cnt_ints, st);
  }
  if (p == 2) {
    // diamonds from the per-edge triangle array: the array's one pass of
    // intersections (k_edge_tri) costs about what one P2 pass costs, and
    // every later diamond or induced-motif query on the graph then only
    // reads it (a few bytes per edge)
    int rc = ensure_sym_aux(g, true, false, st);
    if (rc) return rc;
    const int64_t e0 = g->sym_rp_h[rb], e1 = g->sym_rp_h[re];
    if (e1 <= e0) return JGPM_OK;
    const int64_t quads = (e1 - e0 + 3) / 4;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)8 * c->num_cu, (quads + 255) / 256));
    prof_begin(c, st, "sym_p2");
    hipError_t e = jk::launch_p2_edges(g->d_src, g->sym.ci, g->d_tri, e0, e1, (unsigned long long*)(c->d_out + slot),
                                       grid, st);
    prof_end(c, st);
    return e == hipSuccess ? JGPM_OK : JGPM_E_HIP;
  }
  // The triangle array is shared by later queries; its counter slot is
  // reset per query by ctx_reset_counters before it is used here.
  int rc = ensure_sym_aux(g, query_needs_tri(c, p), p == 3, st);
  if (rc) return rc;
  // the heavy-edge split needs its three edge lists, each as long as the
  // range's edge count up to 2^26 (grown per context, so not while a launch
  // sequence is being captured; edges beyond a full list run in the
  // wave-per-edge launch), and edge indices of a range that fit 32 bits
  const int64_t ne_r = g->sym_rp_h[re] - g->sym_rp_h[rb];
  bool heavy = heavy_split_on(c, p) && ne_r < INT32_MAX;
  const int64_t hwant = std::min<int64_t>(std::max<int64_t>(ne_r, 1), (int64_t)1 << 26);
  if (heavy && c->hcap < hwant) {
    if (c->capturing) {
      heavy = c->d_hlist != nullptr;  // the lists it has
    } else {
      if (c->d_hlist) {
        if (hipStreamSynchronize(st) != hipSuccess) return JGPM_E_HIP;
        (void)hipFree(c->d_hlist);
        c->d_hlist = nullptr;
        c->hcap = 0;
      }
      if (hipMalloc(&c->d_hlist, sizeof(int32_t) * 3 * (size_t)hwant) != hipSuccess) return JGPM_E_NOMEM;
      c->hcap = hwant;
      if (!c->d_hctr && hipMalloc(&c->d_hctr, sizeof(unsigned long long) * 6) != hipSuccess) return JGPM_E_NOMEM;
    }
  }
  const int wpb = jk::sym_waves_per_block();
  const int grid = grid_for(c, 2);
  const size_t stride = (size_t)std::max(1, g->sym_mmin);
  const size_t need = (size_t)grid * wpb * jk::sym_lists_per_wave() * stride;
  if (c->sscr_ints < need) {
    if (c->d_sscr) {
      if (hipStreamSynchronize(st) != hipSuccess) return JGPM_E_HIP;
      (void)hipFree(c->d_sscr);
      c->d_sscr = nullptr;
      c->sscr_ints = 0;
    }
    if (hipMalloc(&c->d_sscr, need * sizeof(int32_t)) != hipSuccess) return JGPM_E_NOMEM;
    c->sscr_ints = need;
  }
  jk::SymArgs a{};
  a.rp = g->sym.rp;
  a.ci = g->sym.ci;
  a.src = g->d_src;
  a.tri = g->d_tri;
  a.nsum = g->d_nsum;
  a.e0 = g->sym_rp_h[rb];
  a.ne = g->sym_rp_h[re] - a.e0;
  if (a.ne <= 0) return JGPM_OK;
  a.qctr = c->d_q64 + (slot % (CTX_SLOTS - 1));
  a.chunk = (p == 2 || p == 3) ? 64 : 8;
  a.scr = c->d_sscr;
  a.scr_stride = (int64_t)stride;
  a.out = c->d_out + slot;
  a.heavy_min = heavy ? c->tune.sym_heavy_min : 0;
  a.heavy_stream = jk::sym_heavy_stream();
  a.nv = g->sym.nv;
  a.hlist = c->d_hlist;
  a.hctr = c->d_hctr;
  a.hcap = (int32_t)c->hcap;
  a.heavy_mid = heavy ? jk::sym_heavy_mid(p) : 0;
  // list counters to zero (a kernel, so the sequence can be captured)
  if (heavy && zero_u64(reinterpret_cast<uint64_t*>(c->d_hctr), 6, st)) return JGPM_E_HIP;
  static const char* const kSymName[13] = {"", "sym_p1", "sym_p2", "sym_p3", "sym_p4", "sym_p5", "sym_p6",
                                           "sym_p7", "sym_p8", "sym_p9", "sym_p10", "sym_p11", "sym_p12"};
  prof_begin(c, st, kSymName[p]);
  // a whole-graph P3 counts each triangle once instead of per directed edge
  hipError_t e = (p == 3 && rb == 0 && re == g->sym.nv) ? jk::launch_p3_triangles(a, grid, st)
                                                        : jk::launch_sym(p, a, grid, st);
  prof_end(c, st);
  if (e != hipSuccess) return JGPM_E_HIP;
  if (heavy) {
    // the edges the wave-per-edge launch listed, in three shapes: |S| up to
    // sym_heavy_mid(p) with a small LDS table and several blocks per CU,
    // up to sym_heavy_nmax() one block per CU, wider ones (graphs of at
    // most 2^20 vertices) with S in the list scratch, one slice per block
    static const char* const kHeavyName[13] = {"", "", "", "", "", "sym_p5_heavy", "", "sym_p7_heavy", "",
                                               "sym_p9_heavy", "", "sym_p11_heavy", "sym_p12_heavy"};
    const int shapes = g->sym.nv <= (1 << 20) ? 3 : 2;
    for (int shape = 0; shape < shapes && e == hipSuccess; ++shape) {
      prof_begin(c, st, kHeavyName[p]);
      e = jk::launch_sym_heavy(p, a, shape, shape ? c->num_cu : 4 * c->num_cu, st);
      prof_end(c, st);
    }
  }
  return e == hipSuccess ? JGPM_OK : JGPM_E_HIP;
}

int enqueue_vertex_sum(jgpm_graph_s* g, int mode, int64_t rb, int64_t re, int slot, hipStream_t st) {
  if (!g->has_sym) return JGPM_E_GRAPHKIND;
  int rc = ensure_sym_aux(g, mode == 1, false, st);
  if (rc) return rc;
  if (re <= rb) return JGPM_OK;
  jgpm_ctx_s* c = exec_ctx(g);
  return jk::launch_vertex_sums(g->sym.rp, g->d_tri, rb, re, mode, (unsigned long long*)(c->d_out + slot),
                                grid_for(c, 4), st) == hipSuccess
             ? JGPM_OK
             : JGPM_E_HIP;
}

// Induced 4-vertex motifs from non-induced counts (exact inverse of the
// conversion matrix): with S = 3-star, T = tailed triangle, C = 4-cycle (P1),
// D = diamond (P2), K = K4 (P13):
//   I_K = K, I_D = D - 6K, I_C = C - D + 3K, I_T = T - 4D + 12K,
//   I_S = S - T + 2D - 4K.
int pattern_slots(int pattern) {
  switch (pattern) {
    case JGPM_I_DIAMOND: return 2;
    case JGPM_I_C4: case JGPM_I_TAILTRI: return 3;
    case JGPM_I_STAR3: return 4;
  }
  return 1;
}

int enqueue_pattern(jgpm_graph_s* g, int pattern, int64_t rb, int64_t re, int s, hipStream_t st) {
  int k = clique_k(pattern);
  if (k) {
    if (!g->has_dag) return JGPM_E_GRAPHKIND;
    return enqueue_clique(g, k, false, rb, re, s, st);
  }
  if (pattern >= JGPM_P1 && pattern <= JGPM_P12) return enqueue_sym(g, pattern, rb, re, s, st);
  if (!g->has_sym) return JGPM_E_GRAPHKIND;
  int rc = JGPM_OK;
  switch (pattern) {
    case JGPM_STAR3: return enqueue_vertex_sum(g, 0, rb, re, s, st);
    case JGPM_TAILTRI: return enqueue_vertex_sum(g, 1, rb, re, s, st);
    case JGPM_I_DIAMOND:
      if ((rc = enqueue_sym(g, JGPM_P2, rb, re, s, st))) return rc;
      return enqueue_clique(g, 4, false, rb, re, s + 1, st);
    case JGPM_I_C4:
      if ((rc = enqueue_sym(g, JGPM_P1, rb, re, s, st))) return rc;
      if ((rc = enqueue_sym(g, JGPM_P2, rb, re, s + 1, st))) return rc;
      return enqueue_clique(g, 4, false, rb, re, s + 2, st);
    case JGPM_I_TAILTRI:
      if ((rc = enqueue_vertex_sum(g, 1, rb, re, s, st))) return rc;
      if ((rc = enqueue_sym(g, JGPM_P2, rb, re, s + 1, st))) return rc;
      return enqueue_clique(g, 4, false, rb, re, s + 2, st);
    case JGPM_I_STAR3:
      if ((rc = enqueue_vertex_sum(g, 0, rb, re, s, st))) return rc;
      if ((rc = enqueue_vertex_sum(g, 1, rb, re, s + 1, st))) return rc;
      if ((rc = enqueue_sym(g, JGPM_P2, rb, re, s + 2, st))) return rc;
      return enqueue_clique(g, 4, false, rb, re, s + 3, st);
  }
  return JGPM_E_PATTERN;
}

int prepare_pattern(jgpm_graph_s* g, int pattern, int64_t rb, int64_t re) {
  jgpm_ctx_s* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->dev) != hipSuccess) return JGPM_E_HIP;
  ExecScope own(c);
  c->nprof = 0;
  int rc = JGPM_OK;
  if (clique_k(pattern)) {
    if (g->has_dag) rc = ensure_huge_ids(g);
  } else if (pattern == JGPM_P1) {
    if (g->has_sym && rb == 0 && re == g->sym.nv) rc = ensure_rank_csr(g, c->st);
  } else if (pattern >= JGPM_P2 && pattern <= JGPM_P12) {
    if (g->has_sym) rc = ensure_sym_aux(g, query_needs_tri(c, pattern), pattern == JGPM_P3, c->st);
  }
  if (hipStreamSynchronize(c->st) != hipSuccess && !rc) rc = JGPM_E_HIP;
  return rc;
}

uint64_t finish_pattern(int pattern, const uint64_t* h) {
  // unsigned wrap-around arithmetic is exact for the non-negative result
  switch (pattern) {
    case JGPM_I_DIAMOND: return h[0] - 6 * h[1];
    case JGPM_I_C4: return h[0] - h[1] + 3 * h[2];
    case JGPM_I_TAILTRI: return h[0] - 4 * h[1] + 12 * h[2];
    case JGPM_I_STAR3: return h[0] - h[1] + 2 * h[2] - 4 * h[3];
  }
  return h[0];
}

}  // namespace jgpm
