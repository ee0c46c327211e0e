// ingest.hip — graph construction on the GPU: an undirected edge list (or a
// seeded R-MAT generated on the device) becomes the resident symmetric CSR and
// degree DAG without a host round trip through sorted lists.
//
// The reference builds, symmetrises and orients on the host
// (common/graph.cc:8-124 loader, :210-401 orientation, rule :229-230).  Here:
//   1. every non-loop edge (u, v) gives two keys (u << B) | v and (v << B) | u
//      (B = bits of the largest id); a loop gives the all-ones key, which
//      decodes to u == v and is dropped below;
//   2. device radix sort of the keys over their 2B significant bits;
//   3. the first copy of every non-loop key is kept: an exclusive scan of the
//      keep flags gives each kept key its CSR slot, so column ids come out
//      ascending within each row and rows in order; degrees are counted and
//      scanned into the row pointers;
//   4. the degree DAG keeps u -> v iff d(v) > d(u) or (d(v) == d(u) and
//      v > u), compacted by a second scan, which keeps every out-list
//      ascending;
//   5. the tier lists, the degree order and the longest-intersection bound
//      complete the graph handle as for a host upload.
// The result is the same graph as jgpm_csr_from_edges / jgpm_gen_rmat
// followed by jgpm_graph_upload(JGPM_SYM); the tests compare them array by
// array.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "jgpm_device.hpp"

namespace {

__global__ void k_rmat_edges(int scale, jgpm::RmatCuts t, uint64_t seed, int64_t m, int32_t* src, int32_t* dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t u, v;
    jgpm::rmat_sample(scale, t, seed, (uint64_t)i, u, v);
    src[i] = (int32_t)u;
    dst[i] = (int32_t)v;
  }
}

__global__ void k_edge_keys(const int32_t* src, const int32_t* dst, int64_t m, int B, uint64_t* keys) {
  const uint64_t loop = (2 * B >= 64) ? ~0ull : ((1ull << (2 * B)) - 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t s = (uint32_t)src[i], d = (uint32_t)dst[i];
    const bool ok = s != d;
    keys[2 * i] = ok ? ((s << B) | d) : loop;
    keys[2 * i + 1] = ok ? ((d << B) | s) : loop;
  }
}

// keep[i] = 1 for the first copy of a key whose endpoints differ
__global__ void k_first_copy(const uint64_t* keys, int64_t n, int B, int64_t* keep) {
  const uint64_t mask = (1ull << B) - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    const bool first = i == 0 || keys[i - 1] != k;
    keep[i] = (first && (k >> B) != (k & mask)) ? 1 : 0;
  }
}

__global__ void k_scatter_sym(const uint64_t* keys, const int64_t* keep, const int64_t* pos, int64_t n, int B,
                              int32_t* col, int32_t* row, unsigned long long* deg) {
  const uint64_t mask = (1ull << B) - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (!keep[i]) continue;
    const uint64_t k = keys[i];
    const int64_t p = pos[i];
    const int32_t u = (int32_t)(k >> B);
    col[p] = (int32_t)(k & mask);
    row[p] = u;
    atomicAdd(&deg[u], 1ull);
  }
}

// orientation rule of the degree DAG, per symmetric edge
__global__ void k_dag_keep(const int32_t* row, const int32_t* col, const int64_t* rp, int64_t nnz, int64_t* keep,
                           int* mmin) {
  int best = 1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = row[e], v = col[e];
    const int64_t du = rp[u + 1] - rp[u], dv = rp[v + 1] - rp[v];
    keep[e] = (dv > du || (dv == du && v > u)) ? 1 : 0;
    best = max(best, (int)min(du, dv));
  }
  atomicMax(mmin, best);
}

__global__ void k_scatter_dag(const int32_t* row, const int32_t* col, const int64_t* keep, const int64_t* pos,
                              int64_t nnz, int32_t* dcol, unsigned long long* deg) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    if (!keep[e]) continue;
    dcol[pos[e]] = col[e];
    atomicAdd(&deg[row[e]], 1ull);
  }
}

__global__ void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Device allocations of one build, released on every exit path; work queued
// on the stream finishes before anything is freed.
struct Scratch {
  hipStream_t st = nullptr;
  std::vector<void*> p;
  template <class T>
  T* get(size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(q);
    return static_cast<T*>(q);
  }
  void drop(void* q) {
    auto it = std::find(p.begin(), p.end(), q);
    if (it != p.end()) {
      (void)hipStreamSynchronize(st);
      (void)hipFree(q);
      p.erase(it);
    }
  }
  ~Scratch() {
    (void)hipStreamSynchronize(st);
    for (void* q : p) (void)hipFree(q);
  }
};

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 16)); }

// exclusive prefix sum of n int64 values (n >= 1) into out
int scan_i64(Scratch& S, const int64_t* in, int64_t* out, int64_t n, hipStream_t st) {
  size_t tb = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, st) != hipSuccess) return JGPM_E_HIP;
  void* tmp = S.get<char>(tb);
  if (!tmp) return JGPM_E_NOMEM;
  if (hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, st) != hipSuccess) return JGPM_E_HIP;
  S.drop(tmp);
  return JGPM_OK;
}

int orient_on_device(jgpm_ctx_s* c, Scratch& S, jgpm_graph_s* g, int32_t* row, int32_t nv, int64_t nnz,
                     bool dag_only);

// Edge list already on the device (d_src, d_dst, m pairs) -> graph handle.
int build_on_device(jgpm_ctx_s* c, Scratch& S, int32_t* d_src, int32_t* d_dst, int64_t m, int32_t nv,
                    jgpm_graph_t* out) {
  hipStream_t st = c->st;
  int B = 1;
  while (B < 31 && (1ll << B) < (int64_t)nv) ++B;
  const int64_t n2 = 2 * m;
  // 1-2: keys and radix sort
  uint64_t* keys = S.get<uint64_t>(n2);
  uint64_t* alt = S.get<uint64_t>(n2);
  if (!keys || !alt) return JGPM_E_NOMEM;
  if (m) hipLaunchKernelGGL(k_edge_keys, dim3(grid_for(m)), dim3(256), 0, st, d_src, d_dst, m, B, keys);
  S.drop(d_src);
  S.drop(d_dst);
  hipcub::DoubleBuffer<uint64_t> db(keys, alt);
  if (n2) {
    size_t tb = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, tb, db, (int)n2, 0, 2 * B, st) != hipSuccess) return JGPM_E_HIP;
    void* tmp = S.get<char>(tb);
    if (!tmp) return JGPM_E_NOMEM;
    if (hipcub::DeviceRadixSort::SortKeys(tmp, tb, db, (int)n2, 0, 2 * B, st) != hipSuccess) return JGPM_E_HIP;
    S.drop(tmp);
  }
  uint64_t* sorted = db.Current();
  // 3: first copies -> symmetric CSR
  int64_t* keep = reinterpret_cast<int64_t*>(db.Alternate());  // the free sort buffer, same size
  int64_t* pos = S.get<int64_t>(n2 + 1);
  unsigned long long* deg = S.get<unsigned long long>((size_t)nv + 1);
  if (!pos || !deg) return JGPM_E_NOMEM;
  int64_t nnz = 0;
  if (n2) {
    hipLaunchKernelGGL(k_first_copy, dim3(grid_for(n2)), dim3(256), 0, st, sorted, n2, B, keep);
    int rc = scan_i64(S, keep, pos, n2, st);
    if (rc) return rc;
    int64_t lastk = 0, lastp = 0;
    if (hipMemcpyAsync(&lastk, keep + n2 - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&lastp, pos + n2 - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return JGPM_E_HIP;
    nnz = lastp + lastk;
  }
  auto* g = new (std::nothrow) jgpm_graph_s();
  if (!g) return JGPM_E_NOMEM;
  g->ctx = c;
  auto fail = [&](int code) {
    jgpm_graph_free(g);
    return code;
  };
  int32_t* row = S.get<int32_t>(nnz);
  if (!row || hipMalloc(&g->sym.ci, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
      hipMalloc(&g->sym.rp, sizeof(int64_t) * ((size_t)nv + 1)) != hipSuccess)
    return fail(JGPM_E_NOMEM);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(nv + 1)), dim3(256), 0, st, reinterpret_cast<int64_t*>(deg),
                     (int64_t)nv + 1, (int64_t)0);
  if (n2)
    hipLaunchKernelGGL(k_scatter_sym, dim3(grid_for(n2)), dim3(256), 0, st, sorted, keep, pos, n2, B, g->sym.ci, row,
                       deg);
  int rc = scan_i64(S, reinterpret_cast<const int64_t*>(deg), g->sym.rp, (int64_t)nv + 1, st);
  if (rc) return fail(rc);
  S.drop(keys);  // both sort buffers (the keys and the keep flags) are done
  S.drop(alt);
  S.drop(pos);
  S.drop(deg);
  rc = orient_on_device(c, S, g, row, nv, nnz, false);
  if (rc) return fail(rc);
  *out = g;
  return JGPM_OK;
}

// Steps 4-5 for a graph whose symmetric CSR (g->sym.rp / ci, nv rows, nnz
// entries, ascending duplicate-free lists) is on the device, with row[e] the
// source of entry e: the degree DAG, compacted in order, then the host
// metadata.  dag_only frees the symmetric arrays afterwards (a DAG-only
// handle, as Graph(prefix, use_dag = true)).
int orient_on_device(jgpm_ctx_s* c, Scratch& S, jgpm_graph_s* g, int32_t* row, int32_t nv, int64_t nnz,
                     bool dag_only) {
  hipStream_t st = c->st;
  auto fail = [&](int code) { return code; };
  int rc = JGPM_OK;
  unsigned long long* deg = S.get<unsigned long long>((size_t)nv + 1);
  if (!deg) return JGPM_E_NOMEM;
  // 4: degree DAG
  int64_t* dkeep = S.get<int64_t>(nnz + 1);
  int64_t* dpos = S.get<int64_t>(nnz + 1);
  int* mmin = S.get<int>(1);
  if (!dkeep || !dpos || !mmin) return fail(JGPM_E_NOMEM);
  if (hipMemsetAsync(mmin, 0, sizeof(int), st) != hipSuccess) return fail(JGPM_E_HIP);
  int64_t dnnz = 0;
  if (nnz) {
    hipLaunchKernelGGL(k_dag_keep, dim3(grid_for(nnz)), dim3(256), 0, st, row, g->sym.ci, g->sym.rp, nnz, dkeep,
                       mmin);
    rc = scan_i64(S, dkeep, dpos, nnz, st);
    if (rc) return fail(rc);
    int64_t lastk = 0, lastp = 0;
    if (hipMemcpyAsync(&lastk, dkeep + nnz - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&lastp, dpos + nnz - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(JGPM_E_HIP);
    dnnz = lastp + lastk;
  }
  if (hipMalloc(&g->dag.ci, sizeof(int32_t) * std::max<int64_t>(dnnz, 1)) != hipSuccess ||
      hipMalloc(&g->dag.rp, sizeof(int64_t) * ((size_t)nv + 1)) != hipSuccess)
    return fail(JGPM_E_NOMEM);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(nv + 1)), dim3(256), 0, st, reinterpret_cast<int64_t*>(deg),
                     (int64_t)nv + 1, (int64_t)0);
  if (nnz)
    hipLaunchKernelGGL(k_scatter_dag, dim3(grid_for(nnz)), dim3(256), 0, st, row, g->sym.ci, dkeep, dpos, nnz,
                       g->dag.ci, deg);
  rc = scan_i64(S, reinterpret_cast<const int64_t*>(deg), g->dag.rp, (int64_t)nv + 1, st);
  if (rc) return fail(rc);
  // 5: host metadata (row pointers, tiers, degree order, bounds)
  g->sym_rp_h.resize((size_t)nv + 1);
  g->dag_rp_h.resize((size_t)nv + 1);
  int mm = 1;
  if (hipMemcpyAsync(g->sym_rp_h.data(), g->sym.rp, 8 * ((size_t)nv + 1), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(g->dag_rp_h.data(), g->dag.rp, 8 * ((size_t)nv + 1), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&mm, mmin, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(JGPM_E_HIP);
  int32_t smax = 0, dmax = 0;
  for (int32_t v = 0; v < nv; ++v) {
    smax = std::max<int32_t>(smax, (int32_t)(g->sym_rp_h[v + 1] - g->sym_rp_h[v]));
    dmax = std::max<int32_t>(dmax, (int32_t)(g->dag_rp_h[v + 1] - g->dag_rp_h[v]));
  }
  g->sym = DevCSR{g->sym.rp, g->sym.ci, nv, nnz, smax};
  g->dag = DevCSR{g->dag.rp, g->dag.ci, nv, dnnz, dmax};
  g->has_sym = g->has_dag = true;
  g->sym_mmin = std::max(1, mm);
  g->bytes = 8 * 2 * ((int64_t)nv + 1) + 4 * (nnz + dnnz);
  rc = jgpm::build_tiers(g, g->dag_rp_h.data(), nv);
  if (!rc) rc = jgpm::set_degree_order(g, g->sym_rp_h, smax);
  if (rc) return fail(rc);
  if (dag_only) {
    if (hipStreamSynchronize(st) != hipSuccess) return JGPM_E_HIP;
    (void)hipFree(g->sym.rp);
    (void)hipFree(g->sym.ci);
    g->sym = DevCSR{};
    g->has_sym = false;
    g->sym_rp_h.clear();
    g->bytes = 8 * ((int64_t)nv + 1) + 4 * dnnz;
  }
  return JGPM_OK;
}

constexpr int64_t kMaxEdges = (int64_t)1 << 30;  // 2m keys must fit the sort's int item count

// ---- rank-relabelled symmetric CSR (4-cycle totals in degree order)
__global__ void k_rank_keys(const int32_t* src, const int32_t* ci, const int32_t* ord, int64_t ne, int B,
                            uint64_t* keys) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x)
    keys[e] = ((uint64_t)(uint32_t)ord[src[e]] << B) | (uint32_t)ord[ci[e]];
}
__global__ void k_rank_deg(const int64_t* rp, const int32_t* ord, int32_t nv, int64_t* deg) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
    deg[ord[v]] = rp[v + 1] - rp[v];
}
__global__ void k_low_bits(const uint64_t* keys, int64_t ne, uint64_t mask, int32_t* ci) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x)
    ci[e] = (int32_t)(keys[e] & mask);
}

}  // namespace

extern "C" int jgpm_graph_from_edges(jgpm_ctx_t ctx, int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                                     jgpm_graph_t* out) {
  if (!ctx || !out || nv < 1 || m < 0 || m >= kMaxEdges || (m > 0 && (!src || !dst))) return JGPM_E_INVALID;
  for (int64_t i = 0; i < m; ++i)
    if (src[i] < 0 || src[i] >= nv || dst[i] < 0 || dst[i] >= nv) return JGPM_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->dev) != hipSuccess) return JGPM_E_HIP;
  Scratch S;
  S.st = ctx->st;
  int32_t* ds = S.get<int32_t>(m);
  int32_t* dd = S.get<int32_t>(m);
  if (!ds || !dd) return JGPM_E_NOMEM;
  if (m && (hipMemcpy(ds, src, 4 * m, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dd, dst, 4 * m, hipMemcpyHostToDevice) != hipSuccess))
    return JGPM_E_HIP;
  return build_on_device(ctx, S, ds, dd, m, nv, out);
}

extern "C" int jgpm_graph_rmat(jgpm_ctx_t ctx, int scale, int edge_factor, double a, double b, double c,
                               uint64_t seed, jgpm_graph_t* out) {
  if (!ctx || !out || scale < 1 || scale > 30 || edge_factor < 1) return JGPM_E_INVALID;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return JGPM_E_INVALID;
  const int32_t nv = (int32_t)(1u << scale);
  const int64_t m = (int64_t)edge_factor << scale;
  if (m >= kMaxEdges) return JGPM_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->dev) != hipSuccess) return JGPM_E_HIP;
  Scratch S;
  S.st = ctx->st;
  int32_t* ds = S.get<int32_t>(m);
  int32_t* dd = S.get<int32_t>(m);
  if (!ds || !dd) return JGPM_E_NOMEM;
  hipLaunchKernelGGL(k_rmat_edges, dim3(grid_for(m)), dim3(256), 0, ctx->st, scale, jgpm::rmat_cuts(a, b, c), seed, m,
                     ds, dd);
  if (hipGetLastError() != hipSuccess) return JGPM_E_HIP;
  return build_on_device(ctx, S, ds, dd, m, nv, out);
}

namespace {
__global__ void k_rows(const int64_t* rp, int32_t nv, int32_t* row) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int ln = threadIdx.x & 63;
  for (int64_t v = wid; v < nv; v += nw)
    for (int64_t e = rp[v] + ln; e < rp[v + 1]; e += 64) row[e] = (int32_t)v;
}
}  // namespace

namespace jgpm {
// A checked symmetric CSR in host memory (a loaded file, possibly mapped)
// -> resident graph, the degree DAG built on the device.
int graph_from_host_sym(jgpm_ctx_s* c, const int64_t* rowptr, const int32_t* colidx, int32_t nv, int64_t nnz,
                        bool dag_only, jgpm_graph_t* out) {
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->dev) != hipSuccess) return JGPM_E_HIP;
  Scratch S;
  S.st = c->st;
  auto* g = new (std::nothrow) jgpm_graph_s();
  if (!g) return JGPM_E_NOMEM;
  g->ctx = c;
  auto fail = [&](int code) {
    jgpm_graph_free(g);
    return code;
  };
  int32_t* row = S.get<int32_t>(nnz);
  if (!row || hipMalloc(&g->sym.ci, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
      hipMalloc(&g->sym.rp, sizeof(int64_t) * ((size_t)nv + 1)) != hipSuccess)
    return fail(JGPM_E_NOMEM);
  if (hipMemcpy(g->sym.rp, rowptr, 8 * ((size_t)nv + 1), hipMemcpyHostToDevice) != hipSuccess ||
      (nnz && hipMemcpy(g->sym.ci, colidx, 4 * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess))
    return fail(JGPM_E_HIP);
  hipLaunchKernelGGL(k_rows, dim3(grid_for((int64_t)nv * 64)), dim3(256), 0, c->st, g->sym.rp, nv, row);
  if (hipGetLastError() != hipSuccess) return fail(JGPM_E_HIP);
  const int rc = orient_on_device(c, S, g, row, nv, nnz, dag_only);
  if (rc) return fail(rc);
  *out = g;
  return JGPM_OK;
}
}  // namespace jgpm

extern "C" int jgpm_graph_download(jgpm_graph_t g, int dag, jgpm_csr_t* out) {
  if (!g || !out) return JGPM_E_INVALID;
  if (!dag && !g->has_sym) return JGPM_E_GRAPHKIND;
  const DevCSR& d = dag ? g->dag : g->sym;
  auto* c = new (std::nothrow) jgpm_csr_s();
  if (!c) return JGPM_E_NOMEM;
  try {
    c->rowptr.resize((size_t)d.nv + 1);
    c->col.resize((size_t)d.nnz);
  } catch (const std::bad_alloc&) {
    delete c;
    return JGPM_E_NOMEM;
  }
  c->nv = d.nv;
  c->nnz = d.nnz;
  c->max_deg = d.max_deg;
  std::lock_guard<std::mutex> lk(g->ctx->mu);
  if (hipSetDevice(g->ctx->dev) != hipSuccess || hipStreamSynchronize(g->ctx->st) != hipSuccess ||
      hipMemcpy(c->rowptr.data(), d.rp, 8 * ((size_t)d.nv + 1), hipMemcpyDeviceToHost) != hipSuccess ||
      (d.nnz && hipMemcpy(c->col.data(), d.ci, 4 * (size_t)d.nnz, hipMemcpyDeviceToHost) != hipSuccess)) {
    delete c;
    return JGPM_E_HIP;
  }
  *out = c;
  return JGPM_OK;
}

namespace jgpm {

// Relabel v -> d_ord[v]: sort the keys (ord[v], ord[w]) of every directed
// edge, so each new list is ascending; the row of ord[v] has d(v) entries.
int ensure_rank_csr(jgpm_graph_s* g, hipStream_t st) {
  if (g->rsym.ci || !g->has_sym) return JGPM_OK;
  jgpm_ctx_s* c = exec_ctx(g);
  if (c->capturing) return JGPM_E_STATE;
  int rc = ensure_sym_aux(g, false, false, st);  // COO sources
  if (rc) return rc;
  const int32_t nv = (int32_t)g->sym.nv;
  const int64_t ne = g->sym.nnz;
  int B = 1;
  while (B < 31 && (1ll << B) < (int64_t)nv) ++B;
  Scratch S;
  S.st = st;
  DevCSR r;
  r.nv = nv;
  r.nnz = ne;
  r.max_deg = g->sym.max_deg;
  if (hipMalloc(&r.rp, sizeof(int64_t) * ((size_t)nv + 1)) != hipSuccess) return JGPM_E_NOMEM;
  if (hipMalloc(&r.ci, sizeof(int32_t) * std::max<int64_t>(ne, 1)) != hipSuccess) {
    (void)hipFree(r.rp);
    return JGPM_E_NOMEM;
  }
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(r.rp);
    (void)hipFree(r.ci);
    return code;
  };
  int64_t* deg = S.get<int64_t>((size_t)nv + 1);
  if (!deg) return fail(JGPM_E_NOMEM);
  hipLaunchKernelGGL(k_rank_deg, dim3(grid_for(nv)), dim3(256), 0, st, g->sym.rp, g->d_ord, nv, deg);
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, st, deg + nv, (int64_t)1, (int64_t)0);
  if ((rc = scan_i64(S, deg, r.rp, (int64_t)nv + 1, st))) return fail(rc);
  if (ne) {
    uint64_t* keys = S.get<uint64_t>(ne);
    uint64_t* alt = S.get<uint64_t>(ne);
    if (!keys || !alt) return fail(JGPM_E_NOMEM);
    hipLaunchKernelGGL(k_rank_keys, dim3(grid_for(ne)), dim3(256), 0, st, g->d_src, g->sym.ci, g->d_ord, ne, B, keys);
    hipcub::DoubleBuffer<uint64_t> db(keys, alt);
    size_t tb = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, tb, db, (int)ne, 0, 2 * B, st) != hipSuccess) return fail(JGPM_E_HIP);
    void* tmp = S.get<char>(tb);
    if (!tmp) return fail(JGPM_E_NOMEM);
    if (hipcub::DeviceRadixSort::SortKeys(tmp, tb, db, (int)ne, 0, 2 * B, st) != hipSuccess) return fail(JGPM_E_HIP);
    hipLaunchKernelGGL(k_low_bits, dim3(grid_for(ne)), dim3(256), 0, st, db.Current(), ne, (1ull << B) - 1, r.ci);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return fail(JGPM_E_HIP);
  g->rsym = r;
  g->bytes += 8 * ((int64_t)nv + 1) + 4 * ne;
  return JGPM_OK;
}

}  // namespace jgpm
