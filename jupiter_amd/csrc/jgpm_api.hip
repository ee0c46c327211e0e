// jgpm_api.hip — the C ABI of include/jgpm.h: device context, resident
// graphs, per-query dispatch to the kernels, status codes.
//
// Replaces the reference driver's per-run setup and dispatch
// (SubgraphSolver, src/jupiter_gpm.cu:27-925): one context per device owns
// one HIP stream; graphs stay resident in HBM between queries; a query picks
// kernels by pattern id instead of by string compare, and returns a status
// instead of asserting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "jgpm_internal.hpp"
#include "kernels.hpp"
#include "jgpm_device.hpp"

using namespace jgpm;

#define HIP_TRY(x)                         \
  do {                                     \
    hipError_t _e = (x);                   \
    if (_e != hipSuccess) return JGPM_E_HIP; \
  } while (0)

extern "C" const char* jgpm_strerror(int s) {
  switch (s) {
    case JGPM_OK: return "ok";
    case JGPM_E_INVALID: return "invalid argument";
    case JGPM_E_PATTERN: return "unknown pattern id";
    case JGPM_E_NOMEM: return "out of memory";
    case JGPM_E_HIP: return "HIP runtime error";
    case JGPM_E_IO: return "graph file missing or malformed";
    case JGPM_E_BUDGET: return "request exceeds the memory budget or a supported size";
    case JGPM_E_GRAPHKIND: return "pattern needs the symmetric graph, which this handle does not hold";
    case JGPM_E_BUSY: return "queue full";
    case JGPM_E_NOTREADY: return "result not ready";
    case JGPM_E_STATE: return "call not valid in the current state";
    case JGPM_E_NODEVICE: return "no HIP device";
  }
  return "unknown status";
}

// ---------------------------------------------------------------------------
// small helper kernels

__global__ void k_fill_u64(uint64_t* p, int n, uint64_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

namespace jgpm {

int zero_u64(uint64_t* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n, 0ull);
  return hipGetLastError() == hipSuccess ? JGPM_OK : JGPM_E_HIP;
}

int ctx_reset_counters(jgpm_ctx_s* c, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(CTX_SLOTS), 0, st, c->d_out, CTX_SLOTS, 0ull);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(CTX_SLOTS), 0, st, c->d_q, CTX_SLOTS, 0u);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(CTX_SLOTS), 0, st, (uint64_t*)c->d_q64, CTX_SLOTS, 0ull);
  return hipGetLastError() == hipSuccess ? JGPM_OK : JGPM_E_HIP;
}

void prof_begin(jgpm_ctx_s* c, hipStream_t st, const char* name) {
  if (!c->prof_on || st != c->st || c->nprof >= jgpm_ctx_s::MAX_PROF) return;
  c->pname[c->nprof] = name;
  (void)hipEventRecord(c->pev[2 * c->nprof], st);
}
void prof_end(jgpm_ctx_s* c, hipStream_t st) {
  if (!c->prof_on || st != c->st || c->nprof >= jgpm_ctx_s::MAX_PROF) return;
  (void)hipEventRecord(c->pev[2 * c->nprof + 1], st);
  c->nprof++;
}

int ensure_scratch(jgpm_ctx_s* c, size_t words) {
  if (c->scr_words >= words) return JGPM_OK;
  if (c->d_scr) {
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipFree(c->d_scr));
    c->d_scr = nullptr;
    c->scr_words = 0;
  }
  if (hipMalloc(&c->d_scr, words * sizeof(uint64_t)) != hipSuccess) return JGPM_E_NOMEM;
  c->scr_words = words;
  return JGPM_OK;
}

static int upload_csr(DevCSR& d, const int64_t* rp, const int32_t* ci, int32_t nv, int64_t nnz,
                      int32_t max_deg) {
  d.nv = nv;
  d.nnz = nnz;
  d.max_deg = max_deg;
  if (hipMalloc(&d.rp, sizeof(int64_t) * (nv + 1)) != hipSuccess) return JGPM_E_NOMEM;
  if (hipMalloc(&d.ci, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess) return JGPM_E_NOMEM;
  HIP_TRY(hipMemcpy(d.rp, rp, sizeof(int64_t) * (nv + 1), hipMemcpyHostToDevice));
  if (nnz) HIP_TRY(hipMemcpy(d.ci, ci, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  return JGPM_OK;
}

static void free_csr(DevCSR& d) {
  if (d.rp) (void)hipFree(d.rp);
  if (d.ci) (void)hipFree(d.ci);
  d = DevCSR{};
}

// Root tier lists of the DAG: vertex ids (ascending) whose out-degree falls
// in each block tier; the wave tier scans the root range directly.
int build_tiers(jgpm_graph_s* g, const int64_t* rp, int32_t nv) {
  for (int t = 0; t < jk::NUM_TIERS; ++t) g->tier_h[t].clear();
  for (int32_t v = 0; v < nv; ++v) {
    int64_t d = rp[v + 1] - rp[v];
    if (d <= jk::tier_max_deg(jk::TIER_WAVE)) continue;
    int t = jk::TIER_LDS2;
    while (t < jk::NUM_TIERS - 1 && d > jk::tier_max_deg(t)) ++t;
    if (d > jk::tier_max_deg(t)) {
      g->huge_h.push_back(v);
      continue;
    }
    g->tier_h[t].push_back(v);
  }
  for (int t = 0; t < jk::NUM_TIERS; ++t) {
    if (g->tier_h[t].empty()) continue;
    size_t n = g->tier_h[t].size();
    if (hipMalloc(&g->tier_d[t], n * sizeof(int32_t)) != hipSuccess) return JGPM_E_NOMEM;
    HIP_TRY(hipMemcpy(g->tier_d[t], g->tier_h[t].data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    g->bytes += (int64_t)(n * sizeof(int32_t));
  }
  return JGPM_OK;
}

// Topological order of the degree DAG used by the clique engine's ranking:
// position of every vertex in (symmetric degree, id) order, by counting sort.
int set_degree_order(jgpm_graph_s* g, const std::vector<int64_t>& sym_rp, int32_t max_deg) {
  const int32_t nv = (int32_t)sym_rp.size() - 1;
  std::vector<int32_t> ord((size_t)std::max(nv, 1));
  std::vector<int64_t> pos((size_t)max_deg + 2, 0);
  for (int32_t v = 0; v < nv; ++v) pos[sym_rp[v + 1] - sym_rp[v] + 1]++;
  for (int32_t k = 0; k <= max_deg; ++k) pos[k + 1] += pos[k];
  for (int32_t v = 0; v < nv; ++v) ord[v] = (int32_t)pos[sym_rp[v + 1] - sym_rp[v]]++;
  if (hipMalloc(&g->d_ord, sizeof(int32_t) * std::max(nv, 1)) != hipSuccess) return JGPM_E_NOMEM;
  if (hipMemcpy(g->d_ord, ord.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice) != hipSuccess) return JGPM_E_HIP;
  g->bytes += 4 * (int64_t)nv;
  return JGPM_OK;
}

static int graph_from_host(jgpm_ctx_s* c, const jgpm_csr_s* sym, const jgpm_csr_s* dag, jgpm_graph_t* out) {
  auto* g = new (std::nothrow) jgpm_graph_s();
  if (!g) return JGPM_E_NOMEM;
  g->ctx = c;
  int st = JGPM_OK;
  HIP_TRY(hipSetDevice(c->dev));
  if (sym) {
    st = upload_csr(g->sym, sym->rowptr.data(), sym->col.data(), sym->nv, sym->nnz, sym->max_deg);
    if (st) { jgpm_graph_free(g); return st; }
    g->has_sym = true;
    g->bytes += 8 * (int64_t)(sym->nv + 1) + 4 * sym->nnz;
    g->sym_rp_h = sym->rowptr;
    // longest intersection list of any edge: max over edges of min(d(u), d(v))
    std::atomic<int32_t> best{1};
    parallel_for(sym->nv, 4096, [&](int64_t b, int64_t e) {
      int32_t m = 1;
      for (int64_t u = b; u < e; ++u) {
        const int32_t du = (int32_t)(sym->rowptr[u + 1] - sym->rowptr[u]);
        for (int64_t k = sym->rowptr[u]; k < sym->rowptr[u + 1]; ++k) {
          const int32_t v = sym->col[k];
          m = std::max(m, std::min(du, (int32_t)(sym->rowptr[v + 1] - sym->rowptr[v])));
        }
      }
      int32_t cur = best.load();
      while (m > cur && !best.compare_exchange_weak(cur, m)) {
      }
    });
    g->sym_mmin = best.load();
  }
  st = upload_csr(g->dag, dag->rowptr.data(), dag->col.data(), dag->nv, dag->nnz, dag->max_deg);
  if (st) { jgpm_graph_free(g); return st; }
  g->has_dag = true;
  g->bytes += 8 * (int64_t)(dag->nv + 1) + 4 * dag->nnz;
  g->dag_rp_h = dag->rowptr;
  st = build_tiers(g, dag->rowptr.data(), dag->nv);
  if (st) { jgpm_graph_free(g); return st; }
  {
    // a topological order of the DAG: (degree, id) when the symmetric degrees
    // are known (the orientation rule), else Kahn's algorithm on the DAG
    if (sym) {
      st = set_degree_order(g, sym->rowptr, sym->max_deg);
      if (st) { jgpm_graph_free(g); return st; }
      *out = g;
      return JGPM_OK;
    }
    std::vector<int32_t> ord((size_t)dag->nv);
    {
      std::vector<int32_t> indeg((size_t)dag->nv, 0), q;
      q.reserve((size_t)dag->nv);
      for (int64_t k = 0; k < dag->nnz; ++k) indeg[dag->col[k]]++;
      for (int32_t v = 0; v < dag->nv; ++v)
        if (!indeg[v]) q.push_back(v);
      for (size_t h = 0; h < q.size(); ++h) {
        int32_t u = q[h];
        ord[u] = (int32_t)h;
        for (int64_t k = dag->rowptr[u]; k < dag->rowptr[u + 1]; ++k)
          if (--indeg[dag->col[k]] == 0) q.push_back(dag->col[k]);
      }
      if ((int64_t)q.size() != dag->nv) { jgpm_graph_free(g); return JGPM_E_INVALID; }  // not acyclic
    }
    if (hipMalloc(&g->d_ord, sizeof(int32_t) * std::max<int64_t>(dag->nv, 1)) != hipSuccess) {
      jgpm_graph_free(g);
      return JGPM_E_NOMEM;
    }
    if (hipMemcpy(g->d_ord, ord.data(), sizeof(int32_t) * dag->nv, hipMemcpyHostToDevice) != hipSuccess) {
      jgpm_graph_free(g);
      return JGPM_E_HIP;
    }
    g->bytes += 4 * (int64_t)dag->nv;
  }
  *out = g;
  return JGPM_OK;
}

}  // namespace jgpm

// ---------------------------------------------------------------------------
static jgpm_tuning default_tuning() {
  jgpm_tuning t{};
  t.small_max_deg = 64;
  t.lds_max_deg = 1024;  // reported only: the tier widths are compiled in
  t.gallop_ratio = 32;
  t.waves_per_block = 4;
  t.blocks_per_cu = 8;
  t.batch_hub_deg = 128;
  t.batch_root_chunk = 16;
  t.batch_edge_chunk = 8;
  t.batch_blocks_per_cu = 0;  // 0: the batched kernel's occupancy
  t.sym_heavy_min = 32;
  return t;
}

extern "C" int jgpm_init(int dev, jgpm_ctx_t* out) {
  if (!out) return JGPM_E_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return JGPM_E_NODEVICE;
  if (dev < 0 || dev >= n) return JGPM_E_INVALID;
  HIP_TRY(hipSetDevice(dev));
  auto* c = new (std::nothrow) jgpm_ctx_s();
  if (!c) return JGPM_E_NOMEM;
  c->dev = dev;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) { delete c; return JGPM_E_HIP; }
  c->num_cu = p.multiProcessorCount;
  c->lds = (int)p.sharedMemPerBlock;
  c->hbm = (int64_t)p.totalGlobalMem;
  snprintf(c->name, sizeof(c->name), "%s", p.gcnArchName);
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return JGPM_E_HIP; }
  if (hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess) {
    delete c;
    return JGPM_E_HIP;
  }
  for (int i = 0; i < 2 * jgpm_ctx_s::MAX_PROF; ++i)
    if (hipEventCreate(&c->pev[i]) != hipSuccess) {
      delete c;
      return JGPM_E_HIP;
    }
  if (hipMalloc(&c->d_out, CTX_SLOTS * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&c->d_q, CTX_SLOTS * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&c->d_q64, CTX_SLOTS * sizeof(unsigned long long)) != hipSuccess) {
    delete c;
    return JGPM_E_NOMEM;
  }
  c->tune = default_tuning();
  // the no-allocation entry point's scratch address on this device
  int st = noalloc_setup(dev, nullptr);
  if (st) {
    (void)jgpm_shutdown(c);
    return st;
  }
  *out = c;
  return JGPM_OK;
}

extern "C" int jgpm_shutdown(jgpm_ctx_t c) {
  if (!c) return JGPM_E_INVALID;
  (void)hipSetDevice(c->dev);
  (void)hipStreamSynchronize(c->st);
  if (c->d_scr) (void)hipFree(c->d_scr);
  if (c->d_out) (void)hipFree(c->d_out);
  if (c->d_q) (void)hipFree(c->d_q);
  if (c->d_q64) (void)hipFree(c->d_q64);
  if (c->d_sscr) (void)hipFree(c->d_sscr);
  if (c->d_hlist) (void)hipFree(c->d_hlist);
  if (c->d_hctr) (void)hipFree(c->d_hctr);
  if (c->d_p1cnt) (void)hipFree(c->d_p1cnt);
  if (c->d_bq) (void)hipFree(c->d_bq);
  if (c->d_bcs) (void)hipFree(c->d_bcs);
  if (c->d_bcnt) (void)hipFree(c->d_bcnt);
  if (c->d_bq32) (void)hipFree(c->d_bq32);
  (void)hipEventDestroy(c->e0);
  (void)hipEventDestroy(c->e1);
  for (int i = 0; i < 2 * jgpm_ctx_s::MAX_PROF; ++i) (void)hipEventDestroy(c->pev[i]);
  (void)hipStreamDestroy(c->st);
  delete c;
  return JGPM_OK;
}

extern "C" int jgpm_device_info(jgpm_ctx_t c, int* num_cu, int* lds_bytes, int64_t* hbm_bytes, char* name,
                                int name_len) {
  if (!c) return JGPM_E_INVALID;
  if (num_cu) *num_cu = c->num_cu;
  if (lds_bytes) *lds_bytes = c->lds;
  if (hbm_bytes) *hbm_bytes = c->hbm;
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", c->name);
  return JGPM_OK;
}

extern "C" int jgpm_get_tuning(jgpm_ctx_t c, jgpm_tuning* t) {
  if (!c || !t) return JGPM_E_INVALID;
  *t = c->tune;
  return JGPM_OK;
}

extern "C" int jgpm_set_tuning(jgpm_ctx_t c, const jgpm_tuning* t) {
  if (!c || !t) return JGPM_E_INVALID;
  // per field: > 0 sets (clamped), 0 keeps, < 0 restores the default
  const jgpm_tuning d = default_tuning();
  auto put = [](int32_t v, int32_t def, int32_t& dst, int32_t lo, int32_t hi) {
    if (v < 0) dst = def;
    else if (v > 0) dst = std::min(std::max(v, lo), hi);
  };
  put(t->small_max_deg, d.small_max_deg, c->tune.small_max_deg, 1, 64);
  put(t->lds_max_deg, d.lds_max_deg, c->tune.lds_max_deg, 1, 1 << 20);
  put(t->gallop_ratio, d.gallop_ratio, c->tune.gallop_ratio, 1, 1 << 20);
  put(t->waves_per_block, d.waves_per_block, c->tune.waves_per_block, 1, 16);
  put(t->blocks_per_cu, d.blocks_per_cu, c->tune.blocks_per_cu, 1, 64);
  put(t->batch_hub_deg, d.batch_hub_deg, c->tune.batch_hub_deg, 64, 128);
  if (c->tune.batch_hub_deg != 64) c->tune.batch_hub_deg = 128;  // the two LDS row tiles the kernel has
  put(t->batch_root_chunk, d.batch_root_chunk, c->tune.batch_root_chunk, 1, 4096);
  put(t->batch_edge_chunk, d.batch_edge_chunk, c->tune.batch_edge_chunk, 1, 4096);
  put(t->batch_blocks_per_cu, d.batch_blocks_per_cu, c->tune.batch_blocks_per_cu, 1, 64);
  put(t->sym_heavy_min, d.sym_heavy_min, c->tune.sym_heavy_min, 1, 1 << 20);
  return JGPM_OK;
}

extern "C" int jgpm_graph_upload(jgpm_ctx_t c, const int64_t* rowptr, const int32_t* colidx, int32_t nv,
                                 int64_t nnz, uint32_t flags, jgpm_graph_t* out) {
  if (!c || !rowptr || !out || nv < 0 || nnz < 0 || (nnz > 0 && !colidx)) return JGPM_E_INVALID;
  if (flags != JGPM_SYM && flags != JGPM_DAG) return JGPM_E_INVALID;
  if (rowptr[0] != 0 || rowptr[nv] != nnz) return JGPM_E_INVALID;
  jgpm_csr_s h;
  h.nv = nv;
  h.nnz = nnz;
  try {
    h.rowptr.assign(rowptr, rowptr + nv + 1);
    h.col.assign(colidx, colidx + nnz);
  } catch (const std::bad_alloc&) {
    return JGPM_E_NOMEM;
  }
  // ascending duplicate-free lists without loops; a symmetric upload must
  // be symmetric, since the library derives the degree DAG from it
  int st = check_csr(h, flags == JGPM_SYM, JGPM_E_INVALID);
  if (st) return st;
  h.max_deg = compute_max_degree(h);
  if (flags == JGPM_DAG) return graph_from_host(c, nullptr, &h, out);
  jgpm_csr_s dag;
  try {
    orient_dag(h, dag);
  } catch (const std::bad_alloc&) {
    return JGPM_E_NOMEM;
  }
  return graph_from_host(c, &h, &dag, out);
}

extern "C" int jgpm_graph_load(jgpm_ctx_t c, const char* prefix, int dag_only, jgpm_graph_t* out) {
  if (!c || !prefix || !out) return JGPM_E_INVALID;
  // the files are mapped, checked where they lie (the file must hold the
  // symmetric graph: the DAG is derived from it), uploaded, and oriented on
  // the device
  MappedCsr m;
  int st = map_csr(prefix, m);
  if (st) return st;
  st = check_csr_arrays(m.nv, m.nnz, m.rp, m.ci, true, JGPM_E_IO);
  if (!st) st = graph_from_host_sym(c, m.rp, m.ci, m.nv, m.nnz, dag_only != 0, out);
  unmap_csr(m);
  return st;
}

extern "C" int jgpm_graph_free(jgpm_graph_t g) {
  if (!g) return JGPM_E_INVALID;
  if (g->nbatch > 0) return JGPM_E_BUSY;  // a prepared batch still launches on it
  (void)hipSetDevice(g->ctx->dev);
  (void)hipStreamSynchronize(g->ctx->st);
  free_csr(g->sym);
  free_csr(g->dag);
  for (int t = 0; t < jk::NUM_TIERS; ++t)
    if (g->tier_d[t]) (void)hipFree(g->tier_d[t]);
  if (g->d_src) (void)hipFree(g->d_src);
  if (g->d_tri) (void)hipFree(g->d_tri);
  if (g->d_nsum) (void)hipFree(g->d_nsum);
  if (g->d_ord) (void)hipFree(g->d_ord);
  if (g->huge_d) (void)hipFree(g->huge_d);
  free_csr(g->rsym);
  delete g;
  return JGPM_OK;
}

extern "C" int jgpm_graph_info(jgpm_graph_t g, int32_t* nv, int64_t* sym_nnz, int64_t* dag_nnz,
                               int32_t* sym_max_deg, int32_t* dag_max_deg) {
  if (!g) return JGPM_E_INVALID;
  if (nv) *nv = g->dag.nv;
  if (sym_nnz) *sym_nnz = g->has_sym ? g->sym.nnz : -1;
  if (dag_nnz) *dag_nnz = g->dag.nnz;
  if (sym_max_deg) *sym_max_deg = g->has_sym ? g->sym.max_deg : -1;
  if (dag_max_deg) *dag_max_deg = g->dag.max_deg;
  return JGPM_OK;
}

extern "C" int jgpm_graph_bytes(jgpm_graph_t g, int64_t* bytes) {
  if (!g || !bytes) return JGPM_E_INVALID;
  *bytes = g->bytes;
  return JGPM_OK;
}

// ---------------------------------------------------------------------------
// Counting

namespace jgpm {

int clique_k(int pattern) {
  switch (pattern) {
    case JGPM_TC: return 3;
    case JGPM_P13: case JGPM_I_K4: return 4;
    case JGPM_P14: return 5;
    case JGPM_P15: return 6;
    case JGPM_P16: return 7;
  }
  return 0;
}

// Enqueue the clique engine for roots [rb, re) accumulating into out_slot.
static const char* const kTierName[jk::NUM_TIERS] = {"clique_wave64", "clique_lds128", "clique_lds256",
                                                     "clique_lds512", "clique_lds1024", "clique_glb2048"};

int glb_tier_grid(const jgpm_ctx_s* c) { return 4 * c->num_cu; }

int ensure_huge_ids(jgpm_graph_s* g) {
  if (g->huge_d || g->huge_h.empty()) return JGPM_OK;
  if (exec_ctx(g)->capturing) return JGPM_E_STATE;
  if (hipMalloc(&g->huge_d, g->huge_h.size() * sizeof(int32_t)) != hipSuccess) return JGPM_E_NOMEM;
  HIP_TRY(hipMemcpy(g->huge_d, g->huge_h.data(), g->huge_h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  g->bytes += (int64_t)(g->huge_h.size() * sizeof(int32_t));
  return JGPM_OK;
}

int enqueue_clique(jgpm_graph_s* g, int k, bool bytes, int64_t rb, int64_t re, int out_slot, hipStream_t st,
                   bool split_slots, int first_tier, uint64_t* out_ptr, unsigned* qctr_base, uint64_t* gscr,
                   size_t gscr_words) {
  jgpm_ctx_s* c = exec_ctx(g);
  if (!g->huge_h.empty()) {
    // roots wider than 2048 (the DAG upload keeps them in huge_h): the
    // level-nest kernel, its lists in the row scratch
    auto lo = std::lower_bound(g->huge_h.begin(), g->huge_h.end(), (int32_t)rb);
    auto hi = std::lower_bound(g->huge_h.begin(), g->huge_h.end(), (int32_t)std::min<int64_t>(re, INT32_MAX));
    const int64_t n = hi - lo;
    if (n > 0) {
      if (bytes) return JGPM_E_BUDGET;  // the work model covers the row tiers only
      const int64_t stride = std::max(1, g->dag.max_deg);
      const int64_t per_block = (int64_t)jk::huge_lists_per_block(k) * stride;  // ints
      const int64_t cap_ints = gscr ? (int64_t)(gscr_words * 2) : ((int64_t)1 << 26);  // 256 MiB of own scratch
      int grid = (int)std::min<int64_t>({n, (int64_t)2 * c->num_cu, cap_ints / per_block});
      if (grid < 1) return JGPM_E_BUDGET;
      int32_t* scr = nullptr;
      if (gscr) {
        scr = reinterpret_cast<int32_t*>(gscr);
      } else {
        int s2 = ensure_scratch(c, (size_t)((grid * per_block + 1) / 2));
        if (s2) return s2;
        scr = reinterpret_cast<int32_t*>(c->d_scr);
      }
      int s3 = ensure_huge_ids(g);
      if (s3) return s3;
      jk::HugeArgs h{};
      h.rp = g->dag.rp;
      h.ci = g->dag.ci;
      h.ids = g->huge_d + (lo - g->huge_h.begin());
      h.ntasks = n;
      h.scr = scr;
      h.stride = stride;
      h.out = out_ptr ? out_ptr : c->d_out + out_slot;
      prof_begin(c, st, "clique_huge");
      hipError_t e = jk::launch_clique_huge(k, h, grid, st);
      prof_end(c, st);
      if (e != hipSuccess) return JGPM_E_HIP;
    }
  }
  const int maxgrid = std::max(1, c->tune.blocks_per_cu) * c->num_cu;
  for (int tier = jk::NUM_TIERS - 1; tier >= first_tier; --tier) {
    jk::CliqueArgs a{};
    a.rp = g->dag.rp;
    a.ci = g->dag.ci;
    a.ord = g->d_ord;
    // roots up to 256 wide search all members of a long list in one pass of
    // their waves, so binary search pays earlier: from a quarter of the ratio
    // up to 128 wide, an eighth at 256 (R-MAT 22, one quarter of the roots:
    // 64-wide tier 3.65 -> 3.1 ms, 128-wide 5.2 -> 4.6 ms, 256-wide 8.4 ->
    // 8.2 ms for P13; the wider tiers measured no gain)
    const int gr = c->tune.gallop_ratio;
    a.ratio = std::max(1, tier == jk::TIER_LDS4 ? gr / 8 : tier < jk::TIER_LDS4 ? gr / 4 : gr);
    a.qctr = qctr_base ? qctr_base + tier : c->d_q + (out_slot * jk::NUM_TIERS + tier) % CTX_SLOTS;
    a.out = out_ptr ? out_ptr : c->d_out + out_slot + (split_slots ? tier : 0);
    a.gscr = nullptr;
    a.dmax = jk::tier_max_deg(tier);
    a.dmin = tier == jk::TIER_WAVE ? (bytes ? 0 : k - 1) : jk::tier_max_deg(tier - 1) + 1;
    int grid;
    if (tier == jk::TIER_WAVE) {
      a.ids = nullptr;
      a.rb = rb;
      a.ntasks = re - rb;
      a.chunk = 16;
      if (a.ntasks <= 0) continue;
      int64_t waves = (a.ntasks + a.chunk - 1) / a.chunk;
      grid = (int)std::min<int64_t>(maxgrid, (waves + 3) / 4);
    } else {
      const auto& v = g->tier_h[tier];
      auto lo = std::lower_bound(v.begin(), v.end(), (int32_t)rb);
      auto hi = std::lower_bound(v.begin(), v.end(), (int32_t)std::min<int64_t>(re, INT32_MAX));
      int64_t n = hi - lo;
      if (n <= 0) continue;
      a.ids = g->tier_d[tier] + (lo - v.begin());
      a.rb = 0;
      a.ntasks = n;
      a.chunk = 1;
      grid = (int)std::min<int64_t>(maxgrid, n);
      if (tier == jk::TIER_LDS16 && !bytes && k >= 4 && !gscr && !c->capturing) {
        // split form: build with 16 waves per root into global row slots,
        // then count; chunks of roots bound the slots to 1 GiB
        const size_t slot = jk::split_slot_words();
        const int64_t per = std::max<int64_t>(1, ((int64_t)1 << 27) / (int64_t)slot);
        const int64_t cap = std::min<int64_t>(n, per);
        int s2 = ensure_scratch(c, slot * (size_t)cap);
        if (s2) return s2;
        for (int64_t b = 0; b < n; b += cap) {
          jk::CliqueArgs s = a;
          s.ids = a.ids + b;
          s.ntasks = std::min<int64_t>(cap, n - b);
          s.gscr = c->d_scr;
          const int sg = (int)std::min<int64_t>(c->num_cu, s.ntasks);
          for (int part = 0; part < 2; ++part) {
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, a.qctr, 1, 0u);
            prof_begin(c, st, part == 0 ? "clique_lds1024_build" : kTierName[tier]);
            hipError_t e = jk::launch_clique_split(k, part == 0, s, sg, st);
            prof_end(c, st);
            if (e != hipSuccess) return JGPM_E_HIP;
          }
        }
        continue;
      }
      if (tier == jk::TIER_GLB32 && !bytes && k >= 5 && !gscr && !c->capturing) {
        // split form: build with 16 waves per root into 512 KiB row slots,
        // then count with several blocks per root sharing its rows through a
        // per-root row counter; chunks of roots bound the slots to 1 GiB
        const size_t slot = jk::glb_slot_words();
        const int64_t per = std::max<int64_t>(1, ((int64_t)1 << 27) / (int64_t)slot);
        const int64_t cap = std::min<int64_t>(n, per);
        int s2 = ensure_scratch(c, slot * (size_t)cap + (size_t)(cap + 8 + 1) / 2);
        if (s2) return s2;
        // 8 per-XCD queue counters, then one row counter per root
        uint32_t* ctrs = reinterpret_cast<uint32_t*>(c->d_scr + slot * (size_t)cap);
        // blocks per root: R-MAT 24 P14, 1 / 2 / 4 / 8 / 16 / 32 blocks: tier
        // (five launch pairs) 1,734 / 1,444 / 1,193 / 1,075 / 1,068 / 1,071 ms
        // for the query (the fused kernel 1,650)
        const int rep = 8;
        for (int64_t b = 0; b < n; b += cap) {
          jk::CliqueArgs s = a;
          s.ids = a.ids + b;
          s.ntasks = std::min<int64_t>(cap, n - b);
          s.gscr = c->d_scr;
          s.rowctr = ctrs + 8;
          s.rep = rep;
          hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((s.ntasks + 8 + 255) / 256)), dim3(256), 0, st, ctrs,
                             (int)(s.ntasks + 8), 0u);
          for (int part = 0; part < 2; ++part) {
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, a.qctr, 1, 0u);
            // resident blocks per CU: build 2 (16 waves, 8 per SIMD); count 1
            // for 5-cliques (the 2x2 tiles' registers), 3 otherwise
            const int per_cu = part == 0 ? 2 : k == 5 ? 1 : 3;
            const int sg = (int)std::min<int64_t>((int64_t)per_cu * c->num_cu, part == 0 ? s.ntasks : s.ntasks * rep);
            prof_begin(c, st, part == 0 ? "clique_glb2048_build" : kTierName[tier]);
            hipError_t e = jk::launch_clique_glb_split(k, part == 0, s, sg, st);
            prof_end(c, st);
            if (e != hipSuccess) return JGPM_E_HIP;
          }
        }
        continue;
      }
      size_t sw = jk::tier_scratch_words(tier);
      if (sw) {
        grid = (int)std::min<int64_t>(grid, glb_tier_grid(c));
        if (gscr) {  // caller-owned scratch (prepared batches): no allocation here
          grid = (int)std::min<int64_t>(grid, (int64_t)(gscr_words / sw));
          if (grid < 1) return JGPM_E_NOMEM;
          a.gscr = gscr;
        } else {
          int s2 = ensure_scratch(c, sw * (size_t)grid);
          if (s2) return s2;
          a.gscr = c->d_scr;
        }
      }
    }
    prof_begin(c, st, kTierName[tier]);
    hipError_t e = jk::launch_clique(k, bytes, tier, a, grid, st);
    prof_end(c, st);
    if (e != hipSuccess) return JGPM_E_HIP;
  }
  return JGPM_OK;
}

}  // namespace jgpm

static int check_range(jgpm_graph_s* g, int64_t& rb, int64_t& re) {
  const int32_t nv = g->dag.nv;
  if (re < 0) re = nv;
  if (rb < 0 || rb > re || re > nv) return JGPM_E_INVALID;
  return JGPM_OK;
}

extern "C" int jgpm_count(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end, uint64_t* count,
                          double* kernel_ms) {
  if (!g || !count) return JGPM_E_INVALID;
  if (pattern < 0 || pattern >= JGPM_NUM_PATTERNS) return JGPM_E_PATTERN;
  int st = check_range(g, root_begin, root_end);
  if (st) return st;
  // The induced motifs combine counts that credit an occurrence to different
  // vertices (P2 to the larger edge end, P13 to the DAG source, the vertex
  // sums to the centre), so only the whole graph gives an exact value.
  if (pattern_slots(pattern) > 1 && (root_begin != 0 || root_end != g->dag.nv)) return JGPM_E_INVALID;
  jgpm_ctx_s* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipEventRecord(c->e0, c->st));
  st = ctx_reset_counters(c, c->st);
  if (st) return st;
  c->nprof = 0;
  st = enqueue_pattern(g, pattern, root_begin, root_end, 0, c->st);
  if (st) {
    (void)hipStreamSynchronize(c->st);
    return st;
  }
  HIP_TRY(hipEventRecord(c->e1, c->st));
  uint64_t h[CTX_SLOTS];
  HIP_TRY(hipMemcpyAsync(h, c->d_out, sizeof(h), hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  *count = finish_pattern(pattern, h);
  if (kernel_ms) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->e0, c->e1));
    *kernel_ms = ms;
  }
  return JGPM_OK;
}

extern "C" int jgpm_work_bytes(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end, uint64_t* bytes) {
  if (!g || !bytes) return JGPM_E_INVALID;
  int k = clique_k(pattern);
  if (!k || pattern == JGPM_I_K4) return JGPM_E_PATTERN;
  int st = check_range(g, root_begin, root_end);
  if (st) return st;
  jgpm_ctx_s* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->dev));
  st = ctx_reset_counters(c, c->st);
  if (st) return st;
  st = enqueue_clique(g, k, true, root_begin, root_end, 0, c->st);
  if (st) {
    (void)hipStreamSynchronize(c->st);
    return st;
  }
  uint64_t h = 0;
  HIP_TRY(hipMemcpyAsync(&h, c->d_out, sizeof(h), hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  *bytes = h;
  return JGPM_OK;
}

extern "C" int jgpm_last_profile(jgpm_ctx_t c, int n, const char** names, double* ms) {
  if (!c || n < 0) return JGPM_E_INVALID;
  std::lock_guard<std::mutex> lk(c->mu);
  for (int i = 0; i < c->nprof && i < n; ++i) {
    float t = 0;
    if (hipEventElapsedTime(&t, c->pev[2 * i], c->pev[2 * i + 1]) != hipSuccess) return JGPM_E_HIP;
    if (names) names[i] = c->pname[i];
    if (ms) ms[i] = t;
  }
  return c->nprof;
}

extern "C" int jgpm_work_bytes_split(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end,
                                     uint64_t* bytes, int n) {
  if (!g || !bytes || n < 0) return JGPM_E_INVALID;
  int k = clique_k(pattern);
  if (!k || pattern == JGPM_I_K4) return JGPM_E_PATTERN;
  int st = check_range(g, root_begin, root_end);
  if (st) return st;
  jgpm_ctx_s* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->dev));
  st = ctx_reset_counters(c, c->st);
  if (st) return st;
  c->nprof = 0;
  st = enqueue_clique(g, k, true, root_begin, root_end, 0, c->st, true);
  if (st) {
    (void)hipStreamSynchronize(c->st);
    return st;
  }
  uint64_t h[CTX_SLOTS];
  HIP_TRY(hipMemcpyAsync(h, c->d_out, sizeof(h), hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  // launches run from the last tier down, skipping empty tiers: report in launch order
  int m = 0;
  for (int tier = jk::NUM_TIERS - 1; tier >= 0 && m < c->nprof; --tier) {
    bool launched = false;
    for (int i = 0; i < c->nprof; ++i) launched |= (c->pname[i] == kTierName[tier]);
    if (!launched) continue;
    if (m < n) bytes[m] = h[tier];
    ++m;
  }
  return m;
}
