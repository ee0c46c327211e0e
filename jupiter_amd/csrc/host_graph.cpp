// Host graph construction: seeded generators, symmetric CSR build,
// degree-DAG orientation and the binary graph file format.
//
// Orientation rule and file layout follow the reference's Graph class
// (common/graph.cc:210-401 for orientation, :26-46 and :727-753 for the
// .meta.txt/.vertex.bin/.edge.bin files).  Everything here is host code: it
// prepares the inputs the device kernels count over.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "jgpm_internal.hpp"

namespace jgpm {

int host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (const char* e = getenv("OMP_NUM_THREADS")) {
    int v = atoi(e);
    if (v > 0) n = (unsigned)v;
  }
  if (n == 0) n = 1;
  if (n > 64) n = 64;
  return (int)n;
}

void parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& f) {
  if (n <= 0) return;
  if (grain < 1) grain = 1;
  int nt = host_threads();
  if (nt == 1 || n <= grain) {
    f(0, n);
    return;
  }
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      int64_t b = next.fetch_add(grain);
      if (b >= n) break;
      f(b, std::min(n, b + grain));
    }
  };
  std::vector<std::thread> ts;
  ts.reserve(nt - 1);
  for (int t = 1; t < nt; ++t) ts.emplace_back(worker);
  worker();
  for (auto& t : ts) t.join();
}

static void exclusive_scan_i32_to_i64(const std::vector<int32_t>& in, std::vector<int64_t>& out) {
  out.resize(in.size() + 1);
  int64_t acc = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    out[i] = acc;
    acc += in[i];
  }
  out[in.size()] = acc;
}

int32_t compute_max_degree(const jgpm_csr_s& c) {
  int32_t md = 0;
  for (int32_t v = 0; v < c.nv; ++v) md = std::max<int32_t>(md, (int32_t)(c.rowptr[v + 1] - c.rowptr[v]));
  return md;
}

// Undirected edge list -> symmetric CSR with ascending, duplicate-free lists
// and no self loops.
int build_sym_csr(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst, jgpm_csr_s& out) {
  if (nv < 0 || m < 0 || (m > 0 && (!src || !dst))) return JGPM_E_INVALID;
  for (int64_t i = 0; i < m; ++i)
    if (src[i] < 0 || src[i] >= nv || dst[i] < 0 || dst[i] >= nv) return JGPM_E_INVALID;
  try {
    std::vector<int32_t> deg(nv, 0);
    parallel_for(m, 1 << 16, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int32_t s = src[i], d = dst[i];
        if (s == d) continue;
        __atomic_fetch_add(&deg[s], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[d], 1, __ATOMIC_RELAXED);
      }
    });
    std::vector<int64_t> off;
    exclusive_scan_i32_to_i64(deg, off);
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    std::vector<int32_t> tmp((size_t)off[nv]);
    parallel_for(m, 1 << 16, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int32_t s = src[i], d = dst[i];
        if (s == d) continue;
        tmp[__atomic_fetch_add(&pos[s], 1, __ATOMIC_RELAXED)] = d;
        tmp[__atomic_fetch_add(&pos[d], 1, __ATOMIC_RELAXED)] = s;
      }
    });
    std::vector<int32_t> ndeg(nv, 0);
    parallel_for(nv, 256, [&](int64_t b, int64_t e) {
      for (int64_t v = b; v < e; ++v) {
        int32_t* p = tmp.data() + off[v];
        int32_t* q = tmp.data() + off[v + 1];
        std::sort(p, q);
        ndeg[v] = (int32_t)(std::unique(p, q) - p);
      }
    });
    out.nv = nv;
    exclusive_scan_i32_to_i64(ndeg, out.rowptr);
    out.nnz = out.rowptr[nv];
    out.col.assign((size_t)out.nnz, 0);
    parallel_for(nv, 1024, [&](int64_t b, int64_t e) {
      for (int64_t v = b; v < e; ++v)
        std::memcpy(out.col.data() + out.rowptr[v], tmp.data() + off[v], sizeof(int32_t) * ndeg[v]);
    });
    out.max_deg = compute_max_degree(out);
  } catch (const std::bad_alloc&) {
    return JGPM_E_NOMEM;
  }
  return JGPM_OK;
}

// Keep u->v iff d(v) > d(u) or (d(v) == d(u) and v > u); source order is kept
// so every out-list stays ascending (common/graph.cc:229-230,377-390).
void orient_dag(const jgpm_csr_s& sym, jgpm_csr_s& dag) {
  const int32_t nv = sym.nv;
  auto deg = [&](int32_t v) { return (int32_t)(sym.rowptr[v + 1] - sym.rowptr[v]); };
  auto keep = [&](int32_t u, int32_t v) {
    int32_t du = deg(u), dv = deg(v);
    return dv > du || (dv == du && v > u);
  };
  std::vector<int32_t> nd(nv, 0);
  parallel_for(nv, 1024, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      int32_t c = 0;
      for (int64_t k = sym.rowptr[u]; k < sym.rowptr[u + 1]; ++k) c += keep((int32_t)u, sym.col[k]);
      nd[u] = c;
    }
  });
  dag.nv = nv;
  exclusive_scan_i32_to_i64(nd, dag.rowptr);
  dag.nnz = dag.rowptr[nv];
  dag.col.assign((size_t)dag.nnz, 0);
  parallel_for(nv, 1024, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      int64_t o = dag.rowptr[u];
      for (int64_t k = sym.rowptr[u]; k < sym.rowptr[u + 1]; ++k)
        if (keep((int32_t)u, sym.col[k])) dag.col[o++] = sym.col[k];
    }
  });
  dag.max_deg = compute_max_degree(dag);
}

// <prefix>.meta.txt holds "nv ne 4 8 1 2 max_degree 0 0 0" whitespace
// separated; .vertex.bin int64[nv+1]; .edge.bin int32[ne].
int write_csr(const jgpm_csr_s& c, const char* prefix) {
  std::string p(prefix);
  std::ofstream fv(p + ".vertex.bin", std::ios::binary);
  std::ofstream fe(p + ".edge.bin", std::ios::binary);
  std::ofstream fm(p + ".meta.txt");
  if (!fv || !fe || !fm) return JGPM_E_IO;
  fv.write(reinterpret_cast<const char*>(c.rowptr.data()), sizeof(int64_t) * (c.nv + 1));
  fe.write(reinterpret_cast<const char*>(c.col.data()), sizeof(int32_t) * c.nnz);
  fm << c.nv << "\n" << c.nnz << "\n4 8 1 2\n" << c.max_deg << "\n0\n0\n0\n";
  if (!fv || !fe || !fm) return JGPM_E_IO;
  return JGPM_OK;
}

int check_csr_arrays(int64_t nv, int64_t nnz, const int64_t* rp, const int32_t* col, bool symmetric, int bad) {
  if (rp[0] != 0 || rp[nv] != nnz) return bad;
  for (int64_t v = 0; v < nv; ++v)
    if (rp[v + 1] < rp[v]) return bad;
  std::atomic<bool> ok{true};
  parallel_for(nv, 1024, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e && ok.load(std::memory_order_relaxed); ++v) {
      const int32_t* l = col + rp[v];
      const int64_t d = rp[v + 1] - rp[v];
      for (int64_t k = 0; k < d; ++k) {
        const int32_t u = l[k];
        if (u < 0 || u >= nv || u == v || (k > 0 && u <= l[k - 1])) { ok = false; break; }
        if (symmetric && u > v) {
          if (!std::binary_search(col + rp[u], col + rp[u + 1], (int32_t)v)) { ok = false; break; }
        }
      }
    }
  });
  if (!ok) return bad;
  if (symmetric) {
    // every u < v entry of a list has its mirror: with the u > v entries found
    // above, symmetry holds iff the two halves have the same size
    std::atomic<int64_t> lo{0}, hi{0};
    parallel_for(nv, 4096, [&](int64_t b, int64_t e) {
      int64_t a = 0, z = 0;
      for (int64_t v = b; v < e; ++v)
        for (int64_t k = rp[v]; k < rp[v + 1]; ++k) (col[k] < v ? a : z) += 1;
      lo += a;
      hi += z;
    });
    if (lo.load() != hi.load()) return bad;
  }
  return JGPM_OK;
}

int check_csr(const jgpm_csr_s& c, bool symmetric, int bad) {
  if ((int64_t)c.rowptr.size() != (int64_t)c.nv + 1 || (int64_t)c.col.size() != c.nnz) return bad;
  return check_csr_arrays(c.nv, c.nnz, c.rowptr.data(), c.col.data(), symmetric, bad);
}

// The reference's files mapped read-only (include/custom_alloc.h maps them
// the same way for out-of-core graphs): no copy into host vectors, the pages
// are read once by the checks and once by the upload.
int map_csr(const char* prefix, MappedCsr& m) {
  std::string p(prefix);
  std::ifstream fm(p + ".meta.txt");
  if (!fm) return JGPM_E_IO;
  int64_t nv = 0, ne = 0;
  int vid_size = 0, eid_size = 0, vl = 0, el = 0, md = 0;
  fm >> nv >> ne >> vid_size >> eid_size >> vl >> el >> md;
  if (!fm || vid_size != 4 || eid_size != 8 || nv <= 0 || nv > INT32_MAX || ne < 0) return JGPM_E_IO;
  auto map_file = [](const std::string& f, size_t want, void*& base, size_t& len) -> int {
    const int fd = open(f.c_str(), O_RDONLY);
    if (fd < 0) return JGPM_E_IO;
    struct stat st;
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < want) {
      close(fd);
      return JGPM_E_IO;
    }
    len = std::max<size_t>(want, 1);
    base = want ? mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
    close(fd);
    if (want && base == MAP_FAILED) {
      base = nullptr;
      return JGPM_E_IO;
    }
    return JGPM_OK;
  };
  int st = map_file(p + ".vertex.bin", sizeof(int64_t) * (size_t)(nv + 1), m.map_v, m.len_v);
  if (!st) st = map_file(p + ".edge.bin", sizeof(int32_t) * (size_t)ne, m.map_e, m.len_e);
  if (st) {
    unmap_csr(m);
    return st;
  }
  m.nv = (int32_t)nv;
  m.nnz = ne;
  m.rp = static_cast<const int64_t*>(m.map_v);
  static const int32_t empty = 0;
  m.ci = ne ? static_cast<const int32_t*>(m.map_e) : &empty;
  return JGPM_OK;
}

void unmap_csr(MappedCsr& m) {
  if (m.map_v) munmap(m.map_v, m.len_v);
  if (m.map_e) munmap(m.map_e, m.len_e);
  m = MappedCsr{};
}

int read_csr(const char* prefix, jgpm_csr_s& c) {
  std::string p(prefix);
  std::ifstream fm(p + ".meta.txt");
  if (!fm) return JGPM_E_IO;
  int64_t nv = 0, ne = 0;
  int vid_size = 0, eid_size = 0, vl = 0, el = 0, md = 0;
  fm >> nv >> ne >> vid_size >> eid_size >> vl >> el >> md;
  if (!fm || vid_size != 4 || eid_size != 8 || nv <= 0 || nv > INT32_MAX || ne < 0) return JGPM_E_IO;
  try {
    c.nv = (int32_t)nv;
    c.nnz = ne;
    c.rowptr.resize((size_t)nv + 1);
    c.col.resize((size_t)ne);
  } catch (const std::bad_alloc&) {
    return JGPM_E_NOMEM;
  }
  std::ifstream fv(p + ".vertex.bin", std::ios::binary);
  std::ifstream fe(p + ".edge.bin", std::ios::binary);
  if (!fv || !fe) return JGPM_E_IO;
  fv.read(reinterpret_cast<char*>(c.rowptr.data()), sizeof(int64_t) * (nv + 1));
  fe.read(reinterpret_cast<char*>(c.col.data()), sizeof(int32_t) * ne);
  if (!fv || !fe) return JGPM_E_IO;
  // a file may hold a DAG (Graph::dump_binary of an oriented graph), so only
  // the list invariants are checked here; loaders that orient check symmetry
  int st = check_csr(c, false, JGPM_E_IO);
  if (st) return st;
  c.max_deg = compute_max_degree(c);
  return JGPM_OK;
}

}  // namespace jgpm

using namespace jgpm;

// ---------------------------------------------------------------------------
// Generators.  Each sample draws from its own splitmix64 stream keyed by
// (seed, sample index), so the edge multiset is independent of thread count.
// Probabilities become 32-bit integer thresholds, so sampling is integer-only
// and bit-reproducible across machines.

RmatCuts jgpm::rmat_cuts(double a, double b, double c) {
  const double T = 4294967296.0;
  return RmatCuts{(uint64_t)std::floor(a * T), (uint64_t)std::floor((a + b) * T), (uint64_t)std::floor((a + b + c) * T)};
}

extern "C" int jgpm_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                             uint64_t seed, jgpm_csr_t* out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1) return JGPM_E_INVALID;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return JGPM_E_INVALID;
  const int32_t nv = (int32_t)(1u << scale);
  const int64_t m = (int64_t)edge_factor << scale;
  const RmatCuts t = rmat_cuts(a, b, c);
  std::vector<int32_t> src, dst;
  try {
    src.resize((size_t)m);
    dst.resize((size_t)m);
  } catch (const std::bad_alloc&) {
    return JGPM_E_NOMEM;
  }
  parallel_for(m, 1 << 15, [&](int64_t bgn, int64_t end) {
    for (int64_t i = bgn; i < end; ++i) {
      uint32_t u, v;
      rmat_sample(scale, t, seed, (uint64_t)i, u, v);
      src[i] = (int32_t)u;
      dst[i] = (int32_t)v;
    }
  });
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  int st = build_sym_csr(nv, m, src.data(), dst.data(), *g);
  if (st) { delete g; return st; }
  *out = g;
  return JGPM_OK;
}

// Vose alias table for sampling index i with probability w[i]/sum(w).
namespace {
struct Alias {
  std::vector<double> prob;
  std::vector<int32_t> alias;
  void build(const std::vector<double>& w) {
    const size_t n = w.size();
    double sum = 0;
    for (double x : w) sum += x;
    prob.assign(n, 0.0);
    alias.assign(n, 0);
    std::vector<double> p(n);
    std::vector<int32_t> small, large;
    for (size_t i = 0; i < n; ++i) {
      p[i] = w[i] * (double)n / sum;
      (p[i] < 1.0 ? small : large).push_back((int32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      int32_t s = small.back(); small.pop_back();
      int32_t l = large.back(); large.pop_back();
      prob[s] = p[s];
      alias[s] = l;
      p[l] = (p[l] + p[s]) - 1.0;
      (p[l] < 1.0 ? small : large).push_back(l);
    }
    for (int32_t i : large) prob[i] = 1.0;
    for (int32_t i : small) prob[i] = 1.0;
  }
  int32_t sample(uint64_t r1, uint64_t r2) const {
    const uint64_t n = prob.size();
    uint64_t i = (uint64_t)(((unsigned __int128)r1 * n) >> 64);
    double u = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);
    return u < prob[i] ? (int32_t)i : alias[i];
  }
};
}  // namespace

extern "C" int jgpm_gen_chung_lu(int32_t n, int64_t target_edges, double gamma, double max_weight,
                                 uint64_t seed, jgpm_csr_t* out) {
  if (!out || n < 2 || target_edges < 1 || gamma <= 1.0 || max_weight <= 0) return JGPM_E_INVALID;
  const double alpha = 1.0 / (gamma - 1.0);
  const double want = 2.0 * (double)target_edges;
  // w_i = max_weight * ((i + i0) / i0)^-alpha; choose i0 so sum w_i = 2E.
  // Fixed-size chunks summed in index order: deterministic for any thread count.
  const int64_t chunk = 1 << 16;
  const int64_t nchunks = (n + chunk - 1) / chunk;
  std::vector<double> part((size_t)nchunks);
  auto total = [&](double i0) {
    parallel_for(nchunks, 1, [&](int64_t b, int64_t e) {
      for (int64_t c = b; c < e; ++c) {
        double s = 0;
        int64_t hi_i = std::min<int64_t>(n, (c + 1) * chunk);
        for (int64_t i = c * chunk; i < hi_i; ++i) s += max_weight * std::pow((i + i0) / i0, -alpha);
        part[c] = s;
      }
    });
    double s = 0;
    for (double x : part) s += x;
    return s;
  };
  double lo = 1e-3, hi = (double)n * 1e3;
  if (total(hi) < want) return JGPM_E_INVALID;  // target unreachable with this max weight
  for (int it = 0; it < 60; ++it) {
    double mid = std::sqrt(lo * hi);
    if (total(mid) < want) lo = mid; else hi = mid;
  }
  const double i0 = hi;
  std::vector<double> w((size_t)n);
  for (int32_t i = 0; i < n; ++i) w[i] = max_weight * std::pow((i + i0) / i0, -alpha);
  Alias al;
  al.build(w);
  std::vector<int32_t> src((size_t)target_edges), dst((size_t)target_edges);
  parallel_for(target_edges, 1 << 15, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      uint64_t st = mix_seed(seed, (uint64_t)i);
      uint64_t r1 = splitmix64(st), r2 = splitmix64(st), r3 = splitmix64(st), r4 = splitmix64(st);
      src[i] = al.sample(r1, r2);
      dst[i] = al.sample(r3, r4);
    }
  });
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  int st = build_sym_csr(n, target_edges, src.data(), dst.data(), *g);
  if (st) { delete g; return st; }
  *out = g;
  return JGPM_OK;
}

extern "C" int jgpm_gen_er(int32_t n, double mean_degree, uint64_t seed, jgpm_csr_t* out) {
  if (!out || n < 2 || mean_degree < 0) return JGPM_E_INVALID;
  const int64_t m = (int64_t)std::llround(mean_degree * (double)n / 2.0);
  std::vector<int32_t> src((size_t)m), dst((size_t)m);
  parallel_for(m, 1 << 15, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      uint64_t st = mix_seed(seed, (uint64_t)i);
      uint64_t r1 = splitmix64(st), r2 = splitmix64(st);
      src[i] = (int32_t)(((unsigned __int128)r1 * (uint64_t)n) >> 64);
      dst[i] = (int32_t)(((unsigned __int128)r2 * (uint64_t)n) >> 64);
    }
  });
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  int st = build_sym_csr(n, m, src.data(), dst.data(), *g);
  if (st) { delete g; return st; }
  *out = g;
  return JGPM_OK;
}

extern "C" int jgpm_gen_ba(int32_t n, int32_t m, uint64_t seed, jgpm_csr_t* out) {
  if (!out || m < 1 || n < m + 1) return JGPM_E_INVALID;
  // Seed graph: a clique on vertices 0..m.  Then each new vertex attaches to
  // m distinct earlier vertices chosen with probability proportional to
  // degree (uniform pick from the endpoint list).
  std::vector<int32_t> src, dst, ends;
  src.reserve((size_t)n * m);
  dst.reserve((size_t)n * m);
  ends.reserve((size_t)2 * n * m + (size_t)m * (m + 1));
  for (int32_t u = 0; u <= m; ++u)
    for (int32_t v = u + 1; v <= m; ++v) {
      src.push_back(u); dst.push_back(v);
      ends.push_back(u); ends.push_back(v);
    }
  uint64_t st = mix_seed(seed, 0x42a);
  std::vector<int32_t> chosen;
  for (int32_t v = m + 1; v < n; ++v) {
    chosen.clear();
    while ((int32_t)chosen.size() < m) {
      uint64_t r = splitmix64(st);
      int32_t t = ends[(size_t)(((unsigned __int128)r * (uint64_t)ends.size()) >> 64)];
      if (std::find(chosen.begin(), chosen.end(), t) == chosen.end()) chosen.push_back(t);
    }
    for (int32_t t : chosen) {
      src.push_back(v); dst.push_back(t);
      ends.push_back(v); ends.push_back(t);
    }
  }
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  int rc = build_sym_csr(n, (int64_t)src.size(), src.data(), dst.data(), *g);
  if (rc) { delete g; return rc; }
  *out = g;
  return JGPM_OK;
}

extern "C" int jgpm_csr_from_edges(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                                   jgpm_csr_t* out) {
  if (!out) return JGPM_E_INVALID;
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  int st = build_sym_csr(nv, m, src, dst, *g);
  if (st) { delete g; return st; }
  *out = g;
  return JGPM_OK;
}

extern "C" int jgpm_csr_orient(jgpm_csr_t sym, jgpm_csr_t* dag) {
  if (!sym || !dag) return JGPM_E_INVALID;
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  try {
    orient_dag(*sym, *g);
  } catch (const std::bad_alloc&) {
    delete g;
    return JGPM_E_NOMEM;
  }
  *dag = g;
  return JGPM_OK;
}

extern "C" int jgpm_csr_info(jgpm_csr_t c, int32_t* nv, int64_t* nnz, int32_t* max_deg) {
  if (!c) return JGPM_E_INVALID;
  if (nv) *nv = c->nv;
  if (nnz) *nnz = c->nnz;
  if (max_deg) *max_deg = c->max_deg;
  return JGPM_OK;
}

extern "C" int jgpm_csr_data(jgpm_csr_t c, const int64_t** rowptr, const int32_t** colidx) {
  if (!c) return JGPM_E_INVALID;
  if (rowptr) *rowptr = c->rowptr.data();
  if (colidx) *colidx = c->col.data();
  return JGPM_OK;
}

extern "C" int jgpm_csr_write(jgpm_csr_t c, const char* prefix) {
  if (!c || !prefix) return JGPM_E_INVALID;
  return write_csr(*c, prefix);
}

extern "C" int jgpm_csr_read(const char* prefix, jgpm_csr_t* out) {
  if (!prefix || !out) return JGPM_E_INVALID;
  auto* g = new (std::nothrow) jgpm_csr_s();
  if (!g) return JGPM_E_NOMEM;
  int st = read_csr(prefix, *g);
  if (st) { delete g; return st; }
  *out = g;
  return JGPM_OK;
}

extern "C" int jgpm_csr_free(jgpm_csr_t c) {
  delete c;
  return JGPM_OK;
}
