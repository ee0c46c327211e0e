// oracle.cc — CPU restatement of the reference's pattern-counting loop nests.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the timed
// CPU baseline ("port" kind).  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product (libjgpm) never does.
//
// What it restates (clean-room; no reference code is copied):
//  * the message-passing kernels P1..P16 of
//    src/message_passing_gpu_kernels.cuh (line ranges on each function), one
//    directed COO edge (v0, v1) at a time, v0 in the requested root range,
//    exactly the loop nest, order filters and exclusion tests of the kernel;
//  * the triangle count of P0_kernel (:64-102) evaluated on the degree DAG,
//    which equals total[1] of pattern_task_generate (src/gpu_kernels_codegen.cuh:77);
//  * the warp set primitives as sequential sorted-list operations:
//    intersect (include/set_intersect.cuh:109-115), intersect with upper
//    bound (:189-195), intersect_num (:353-360), intersect_num with upper
//    bound (:429-437), intersect_num with an ancestor list (:474-507) and
//    count_smaller (include/operations.cuh:80-124, exact because every caller
//    passes a bound that is in the list);
//  * the CPU merge of VertexSet (include/VertexSet.h:53-122) with the
//    merge/galloping switch of common/intersect.cc:11,54 (size ratio > 32
//    for counting, > 50 for materialising); galloping is the textbook
//    exponential search followed by binary search;
//  * Graph::orientation (common/graph.cc:210-401).
//
// Two evaluation modes give the same integer: mode 0 runs the innermost
// count++ loops literally (cost proportional to the count); mode 1 replaces
// each innermost "for x in L: if x not in E: count++" by |L| - |L ∩ E| for
// the distinct excluded vertices E.  Tests check mode 0 == mode 1 on small
// graphs and both against brute-force subgraph counts.
//
// Parity pins (tests/test_oracle.py): brute-force injective-map counts
// divided by |Aut(H)| on random graphs up to 12 vertices; the closed form
// count(H, K_n) = C(n,k) k!/|Aut(H)|; closed forms for TC (trace(A^3)/6), P1
// (1/2 sum_{u<w} C(codeg,2)) and P2 (sum_e C(t_e,2)); relabelling invariance
// and quarter-sum invariance.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

typedef int32_t vid;
typedef uint64_t acc_t;

struct G {
  const int64_t* rp;
  const int32_t* ci;
  int32_t nv;
  const vid* N(vid v) const { return ci + rp[v]; }
  vid d(vid v) const { return (vid)(rp[v + 1] - rp[v]); }
};

// first index i in [b, e) with a[i] >= t (exponential then binary search)
static inline int64_t gallop(const vid* a, int64_t b, int64_t e, vid t) {
  if (b >= e || a[b] >= t) return b;
  int64_t step = 1, lo = b, hi = b + 1;
  while (hi < e && a[hi] < t) {
    lo = hi;
    step <<= 1;
    hi = b + step;
  }
  if (hi > e) hi = e;
  // a[lo] < t, answer in (lo, hi]
  int64_t L = lo + 1, R = hi;
  while (L < R) {
    int64_t m = (L + R) >> 1;
    if (a[m] < t) L = m + 1; else R = m;
  }
  return L;
}

// a ∩ b, ascending, optional strict upper bound.
static vid isect(const vid* a, vid na, const vid* b, vid nb, vid* out, vid upper = INT32_MAX) {
  vid n = 0;
  if (na == 0 || nb == 0) return 0;
  if (na / 50 > nb || nb / 50 > na) {
    const vid *s = a, *l = b;
    vid ns = na, nl = nb;
    if (na > nb) { s = b; l = a; ns = nb; nl = na; }
    int64_t p = 0;
    for (vid i = 0; i < ns; ++i) {
      vid x = s[i];
      if (x >= upper) break;
      p = gallop(l, p, nl, x);
      if (p >= nl) break;
      if (l[p] == x) out[n++] = x;
    }
    return n;
  }
  vid i = 0, j = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x >= upper || y >= upper) break;
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x == y) out[n++] = x;
  }
  return n;
}

static acc_t isect_num(const vid* a, vid na, const vid* b, vid nb, vid upper = INT32_MAX) {
  acc_t n = 0;
  if (na == 0 || nb == 0) return 0;
  if (na / 32 > nb || nb / 32 > na) {
    const vid *s = a, *l = b;
    vid ns = na, nl = nb;
    if (na > nb) { s = b; l = a; ns = nb; nl = na; }
    int64_t p = 0;
    for (vid i = 0; i < ns; ++i) {
      vid x = s[i];
      if (x >= upper) break;
      p = gallop(l, p, nl, x);
      if (p >= nl) break;
      n += (l[p] == x);
    }
    return n;
  }
  vid i = 0, j = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x >= upper || y >= upper) break;
    if (x <= y) ++i;
    if (y <= x) ++j;
    n += (x == y);
  }
  return n;
}

// |{x in a ∩ b : x not in anc[0..k)}|
static acc_t isect_num_excl(const vid* a, vid na, const vid* b, vid nb, const vid* anc, int k) {
  acc_t n = 0;
  vid i = 0, j = 0;
  while (i < na && j < nb) {
    vid x = a[i], y = b[j];
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x == y) {
      bool ok = true;
      for (int t = 0; t < k; ++t) ok &= (x != anc[t]);
      n += ok;
    }
  }
  return n;
}

// number of elements of sorted a that are < bound
static acc_t count_smaller(vid bound, const vid* a, vid na) {
  return (acc_t)(std::lower_bound(a, a + na, bound) - a);
}

static inline bool in_sorted(const vid* a, vid na, vid x) {
  return std::binary_search(a, a + na, x);
}

// |L| minus the number of distinct excluded values present in L.
static inline acc_t count_excluding(const vid* L, vid nL, std::initializer_list<vid> ex) {
  acc_t c = (acc_t)nL;
  vid seen[8];
  int ns = 0;
  for (vid e : ex) {
    bool dup = false;
    for (int t = 0; t < ns; ++t) dup |= (seen[t] == e);
    if (dup) continue;
    seen[ns++] = e;
    if (in_sorted(L, nL, e)) --c;
  }
  return c;
}

struct Scratch {
  std::vector<vid> s, l1, l2, l3, l4, l5;
  explicit Scratch(vid md) : s(md + 1), l1(md + 1), l2(md + 1), l3(md + 1), l4(md + 1), l5(md + 1) {}
};

// One directed edge (v0, v1) of pattern p; returns its contribution.
static acc_t edge_count(const G& g, int p, vid v0, vid v1, int mode, Scratch& w) {
  acc_t count = 0;
  const vid* N0 = g.N(v0);
  const vid d0 = g.d(v0);
  const vid* N1 = g.N(v1);
  const vid d1 = g.d(v1);
  vid* S = w.s.data();
  switch (p) {
    case 0: {  // TC on the DAG (P0_kernel nest)
      return isect_num(N0, d0, N1, d1);
    }
    case 1: {  // P1 :104-186
      if (v1 >= v0) return 0;
      for (vid i = 0; i < d0; ++i) {
        vid v2 = N0[i];
        if (v2 >= v1) break;
        count += isect_num(N1, d1, g.N(v2), g.d(v2), v0);
      }
      return count;
    }
    case 2: {  // P2 :188-252
      if (v1 >= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      for (vid i = 0; i < ns; ++i) count += count_smaller(S[i], S, ns);
      return count;
    }
    case 3: {  // P3 :254-334
      vid ns = isect(N0, d0, N1, d1, S);
      for (vid i3 = 0; i3 < d0; ++i3) {
        vid v3 = N0[i3];
        if (v3 == v1) continue;
        const vid* N3 = g.N(v3);
        vid d3 = g.d(v3);
        for (vid i = 0; i < ns; ++i) {
          vid v2 = S[i];
          if (v2 <= v1 || v2 == v3 || v2 == v0) continue;
          if (mode == 0) {
            for (vid k = 0; k < d3; ++k) {
              vid v4 = N3[k];
              if (v4 == v0 || v4 == v1 || v4 == v3 || v4 == v2) continue;
              ++count;
            }
          } else {
            count += count_excluding(N3, d3, {v0, v1, v3, v2});
          }
        }
      }
      return count;
    }
    case 4: {  // P4 :336-420
      if (v1 >= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L = w.l1.data();
      for (vid i3 = 0; i3 < d1; ++i3) {
        vid v3 = N1[i3];
        if (v3 == v0) continue;
        vid nl = isect(N0, d0, g.N(v3), g.d(v3), L);
        for (vid i = 0; i < ns; ++i) {
          vid v2 = S[i];
          if (v3 == v2) continue;
          if (mode == 0) {
            for (vid k = 0; k < nl; ++k) {
              vid x = L[k];
              if (x == v1 || x == v2) continue;
              ++count;
            }
          } else {
            count += count_excluding(L, nl, {v1, v2});
          }
        }
      }
      return count;
    }
    case 5: {  // P5 :422-502
      if (v1 >= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L = w.l1.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        vid nl = isect(S, ns, g.N(v2), g.d(v2), L, v2);
        for (vid i = 0; i < ns; ++i) {
          vid v4 = S[i];
          if (v2 == v4) continue;
          if (mode == 0) {
            for (vid k = 0; k < nl; ++k) {
              if (L[k] == v4) continue;
              ++count;
            }
          } else {
            count += count_excluding(L, nl, {v4});
          }
        }
      }
      return count;
    }
    case 6: {  // P6 :504-576
      if (v1 >= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      vid anc[2] = {v0, v1};
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        for (vid i3 = 0; i3 < i2; ++i3) {
          vid v3 = S[i3];
          count += isect_num_excl(g.N(v2), g.d(v2), g.N(v3), g.d(v3), anc, 2);
        }
      }
      return count;
    }
    case 7: {  // P7 :578-659
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L = w.l1.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        if (v2 <= v1) continue;
        vid nl = isect(N0, d0, g.N(v2), g.d(v2), L);
        for (vid i = 0; i < nl; ++i) {
          vid v4 = L[i];
          if (v1 == v4) continue;
          if (mode == 0) {
            for (vid k = 0; k < ns; ++k) {
              vid x = S[k];
              if (x == v4 || x == v2) continue;
              ++count;
            }
          } else {
            count += count_excluding(S, ns, {v4, v2});
          }
        }
      }
      return count;
    }
    case 8: {  // P8 :661-745
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L = w.l1.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        if (v2 >= v1) continue;
        vid nl = isect(N0, d0, g.N(v2), g.d(v2), L);
        for (vid i = 0; i < nl; ++i) {
          vid v3 = L[i];
          if (v3 >= v1) continue;
          count += isect_num(S, ns, g.N(v3), g.d(v3), v2);
        }
      }
      return count;
    }
    case 9: {  // P9 :747-825
      if (v1 <= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L = w.l1.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        if (v2 <= v1) continue;
        vid nl = isect(S, ns, g.N(v2), g.d(v2), L);
        for (vid i = 0; i < nl; ++i) count += count_smaller(L[i], L, nl);
      }
      return count;
    }
    case 10: {  // P10 :827-918
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L1 = w.l1.data();
      vid* L2 = w.l2.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        vid n1 = isect(N0, d0, g.N(v2), g.d(v2), L1);
        for (vid i = 0; i < n1; ++i) {
          vid v3 = L1[i];
          if (v3 <= v1) continue;
          vid n2 = isect(N0, d0, g.N(v3), g.d(v3), L2);
          for (vid i4 = 0; i4 < ns; ++i4) {
            vid v4 = S[i4];
            if (v4 == v2 || v4 == v1 || v4 == v3) continue;
            if (mode == 0) {
              for (vid k = 0; k < n2; ++k) {
                vid x = L2[k];
                if (x == v4 || x == v1 || x == v2 || x == v3) continue;
                ++count;
              }
            } else {
              count += count_excluding(L2, n2, {v4, v1, v2, v3});
            }
          }
        }
      }
      return count;
    }
    case 11: {  // P11 :920-1021
      if (v1 <= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L4 = w.l1.data();
      vid* L5 = w.l2.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        if (v2 <= v1) continue;
        vid n4 = isect(N0, d0, g.N(v2), g.d(v2), L4);
        vid n5 = isect(N1, d1, g.N(v2), g.d(v2), L5);
        for (vid i3 = 0; i3 < ns; ++i3) {
          vid v3 = S[i3];
          if (v3 == v2 || v3 == v1 || v3 == v0) continue;
          for (vid i4 = 0; i4 < n4; ++i4) {
            vid v4 = L4[i4];
            if (v4 == v3 || v4 == v2 || v4 == v1) continue;
            if (mode == 0) {
              for (vid k = 0; k < n5; ++k) {
                vid x = L5[k];
                if (x == v4 || x == v1 || x == v2 || x == v3 || x == v0) continue;
                ++count;
              }
            } else {
              count += count_excluding(L5, n5, {v4, v1, v2, v3, v0});
            }
          }
        }
      }
      return count;
    }
    case 12: {  // P12 :1023-1109
      if (v1 <= v0) return 0;
      vid ns = isect(N0, d0, N1, d1, S);
      vid* L = w.l1.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        vid nl = isect(S, ns, g.N(v2), g.d(v2), L);
        for (vid i3 = 0; i3 < nl; ++i3) {
          vid v3 = L[i3];
          for (vid i4 = i3 + 1; i4 < nl; ++i4) {
            vid v4 = L[i4];
            if (mode == 0) {
              for (vid k = 0; k < ns; ++k) {
                vid x = S[k];
                if (x == v4 || x == v1 || x == v2 || x == v3 || x == v0) continue;
                ++count;
              }
            } else {
              count += count_excluding(S, ns, {v4, v1, v2, v3, v0});
            }
          }
        }
      }
      return count;
    }
    case 13: {  // P13 :1208-1286 (on the DAG)
      vid ns = isect(N0, d0, N1, d1, S);
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        count += isect_num(S, ns, g.N(v2), g.d(v2));
      }
      return count;
    }
    case 14: {  // P14 :1288-1356
      vid ns = isect(N0, d0, N1, d1, S);
      vid* T = w.l1.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        vid nt = isect(S, ns, g.N(v2), g.d(v2), T);
        for (vid i3 = 0; i3 < nt; ++i3) {
          vid v3 = T[i3];
          count += isect_num(T, nt, g.N(v3), g.d(v3));
        }
      }
      return count;
    }
    case 15: {  // P15 :1358-1440
      vid ns = isect(N0, d0, N1, d1, S);
      vid* T = w.l1.data();
      vid* U = w.l2.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        vid nt = isect(S, ns, g.N(v2), g.d(v2), T);
        for (vid i3 = 0; i3 < nt; ++i3) {
          vid v3 = T[i3];
          vid nu = isect(T, nt, g.N(v3), g.d(v3), U);
          for (vid i4 = 0; i4 < nu; ++i4) {
            vid v4 = U[i4];
            count += isect_num(U, nu, g.N(v4), g.d(v4));
          }
        }
      }
      return count;
    }
    case 16: {  // P16 :1442-1537
      vid ns = isect(N0, d0, N1, d1, S);
      vid* T = w.l1.data();
      vid* U = w.l2.data();
      vid* V = w.l3.data();
      for (vid i2 = 0; i2 < ns; ++i2) {
        vid v2 = S[i2];
        vid nt = isect(S, ns, g.N(v2), g.d(v2), T);
        for (vid i3 = 0; i3 < nt; ++i3) {
          vid v3 = T[i3];
          vid nu = isect(T, nt, g.N(v3), g.d(v3), U);
          for (vid i4 = 0; i4 < nu; ++i4) {
            vid v4 = U[i4];
            vid nw = isect(U, nu, g.N(v4), g.d(v4), V);
            for (vid i5 = 0; i5 < nw; ++i5) {
              vid v5 = V[i5];
              count += isect_num(V, nw, g.N(v5), g.d(v5));
            }
          }
        }
      }
      return count;
    }
  }
  return 0;
}

// Work model: every neighbour-list fetch of the canonical nest costs
// 16 + 4 d(v) bytes (two int64 row pointers plus the int32 list); the root
// list once per root vertex; intermediate lists are on chip.
static acc_t edge_bytes(const G& g, int p, vid v0, vid v1, Scratch& w) {
  auto f = [&](vid v) -> acc_t { return 16 + 4 * (acc_t)g.d(v); };
  acc_t b = f(v1);
  if (p == 0) return b;
  vid* S = w.s.data();
  vid ns = isect(g.N(v0), g.d(v0), g.N(v1), g.d(v1), S);
  vid* T = w.l1.data();
  vid* U = w.l2.data();
  vid* V = w.l3.data();
  for (vid i2 = 0; i2 < ns; ++i2) {
    vid v2 = S[i2];
    b += f(v2);
    if (p == 13) continue;
    vid nt = isect(S, ns, g.N(v2), g.d(v2), T);
    for (vid i3 = 0; i3 < nt; ++i3) {
      vid v3 = T[i3];
      b += f(v3);
      if (p == 14) continue;
      vid nu = isect(T, nt, g.N(v3), g.d(v3), U);
      for (vid i4 = 0; i4 < nu; ++i4) {
        vid v4 = U[i4];
        b += f(v4);
        if (p == 15) continue;
        vid nw = isect(U, nu, g.N(v4), g.d(v4), V);
        for (vid i5 = 0; i5 < nw; ++i5) b += f(V[i5]);
      }
    }
  }
  return b;
}

}  // namespace

extern "C" {

// Count pattern p (0 = TC on a DAG, 1..16 = P1..P16; P1..P12 expect the
// symmetric CSR, 0 and 13..16 the degree DAG) over roots [rb, re).
int oracle_count(const int64_t* rowptr, const int32_t* col, int32_t nv, int p, int64_t rb,
                 int64_t re, int mode, int nthreads, uint64_t* out) {
  if (!rowptr || !out || p < 0 || p > 16 || rb < 0 || re > nv || rb > re) return -1;
  G g{rowptr, col, nv};
  vid md = 0;
  for (vid v = 0; v < nv; ++v) md = std::max(md, g.d(v));
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  acc_t total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
  {
    Scratch w(md);
#pragma omp for schedule(dynamic, 16)
    for (int64_t v0 = rb; v0 < re; ++v0) {
      for (int64_t e = rowptr[v0]; e < rowptr[v0 + 1]; ++e)
        total += edge_count(g, p, (vid)v0, col[e], mode, w);
    }
  }
  *out = total;
  return 0;
}

// Same as oracle_count but over an explicit list of roots (a bounded
// sample for the CPU baseline timing).
int oracle_count_roots(const int64_t* rowptr, const int32_t* col, int32_t nv, int p,
                       const int32_t* roots, int64_t nroots, int mode, int nthreads,
                       uint64_t* out) {
  if (!rowptr || !out || !roots || p < 0 || p > 16) return -1;
  G g{rowptr, col, nv};
  vid md = 0;
  for (vid v = 0; v < nv; ++v) md = std::max(md, g.d(v));
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  acc_t total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
  {
    Scratch w(md);
#pragma omp for schedule(dynamic, 16)
    for (int64_t r = 0; r < nroots; ++r) {
      vid v0 = roots[r];
      for (int64_t e = rowptr[v0]; e < rowptr[v0 + 1]; ++e) total += edge_count(g, p, v0, col[e], mode, w);
    }
  }
  *out = total;
  return 0;
}

int oracle_work_bytes(const int64_t* rowptr, const int32_t* col, int32_t nv, int p, int64_t rb,
                      int64_t re, int nthreads, uint64_t* out) {
  if (!rowptr || !out || !(p == 0 || (p >= 13 && p <= 16)) || rb < 0 || re > nv || rb > re) return -1;
  G g{rowptr, col, nv};
  vid md = 0;
  for (vid v = 0; v < nv; ++v) md = std::max(md, g.d(v));
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  acc_t total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
  {
    Scratch w(md);
#pragma omp for schedule(dynamic, 16)
    for (int64_t v0 = rb; v0 < re; ++v0) {
      total += 16 + 4 * (acc_t)g.d((vid)v0);
      for (int64_t e = rowptr[v0]; e < rowptr[v0 + 1]; ++e) total += edge_bytes(g, p, (vid)v0, col[e], w);
    }
  }
  *out = total;
  return 0;
}

// Non-induced 3-stars (mode 0: sum_v C(d_v, 3)) and tailed triangles
// (mode 1: sum_v t_v (d_v - 2), t_v the triangles at v) over the centres
// v in [rb, re) of a symmetric CSR: the extension ids of the induced-motif
// workload (textbook closed forms); each triangle
// u < w < x is listed once by a merge of N(u) and N(w) above w.
int oracle_vertex_sums(const int64_t* rowptr, const int32_t* col, int32_t nv, int mode, int64_t rb, int64_t re,
                       uint64_t* out) {
  if (!rowptr || !out || rb < 0 || re > nv || rb > re || (mode != 0 && mode != 1)) return -1;
  G g{rowptr, col, nv};
  std::vector<acc_t> t(mode == 1 ? (size_t)nv : 0, 0);
  if (mode == 1) {
    for (vid u = 0; u < nv; ++u) {
      const vid* Nu = g.N(u);
      const vid du = g.d(u);
      for (vid i = 0; i < du; ++i) {
        const vid w = Nu[i];
        if (w <= u) continue;
        // x in N(u) ∩ N(w), x > w
        const vid* Nw = g.N(w);
        vid a = 0, b = 0;
        const vid dw = g.d(w);
        while (a < du && b < dw) {
          const vid x = Nu[a], y = Nw[b];
          if (x <= y) ++a;
          if (y <= x) ++b;
          if (x == y && x > w) {
            t[u]++;
            t[w]++;
            t[x]++;
          }
        }
      }
    }
  }
  acc_t total = 0;
  for (int64_t v = rb; v < re; ++v) {
    const acc_t d = (acc_t)g.d((vid)v);
    if (mode == 0) total += d >= 3 ? d * (d - 1) * (d - 2) / 6 : 0;
    else total += d >= 2 ? t[v] * (d - 2) : 0;
  }
  *out = total;
  return 0;
}

// Algorithmic bytes of the bitmap clique engine over the roots [rb, re) of a
// DAG: every root reads its own list and each member's list once, at 16 + 4 d
// bytes per list (two int64 row pointers plus the int32 ids).
int oracle_design_bytes(const int64_t* rowptr, const int32_t* col, int32_t nv, int64_t rb, int64_t re,
                        uint64_t* out) {
  if (!rowptr || !out || rb < 0 || re > nv || rb > re) return -1;
  acc_t total = 0;
  for (int64_t v = rb; v < re; ++v) {
    total += 16 + 4 * (acc_t)(rowptr[v + 1] - rowptr[v]);
    for (int64_t e = rowptr[v]; e < rowptr[v + 1]; ++e) total += 16 + 4 * (acc_t)(rowptr[col[e] + 1] - rowptr[col[e]]);
  }
  *out = total;
  return 0;
}

// Graph::orientation restated: returns the DAG nnz when out_col is NULL,
// otherwise fills out_rowptr[nv+1] and out_col[nnz].
int64_t oracle_orient(const int64_t* rowptr, const int32_t* col, int32_t nv, int64_t* out_rowptr,
                      int32_t* out_col) {
  std::vector<vid> deg(nv), nd(nv, 0);
  for (vid v = 0; v < nv; ++v) deg[v] = (vid)(rowptr[v + 1] - rowptr[v]);
  for (vid s = 0; s < nv; ++s)
    for (int64_t k = rowptr[s]; k < rowptr[s + 1]; ++k) {
      vid d = col[k];
      if (deg[d] > deg[s] || (deg[d] == deg[s] && d > s)) nd[s]++;
    }
  int64_t nnz = 0;
  for (vid v = 0; v < nv; ++v) nnz += nd[v];
  if (!out_col || !out_rowptr) return nnz;
  out_rowptr[0] = 0;
  for (vid v = 0; v < nv; ++v) out_rowptr[v + 1] = out_rowptr[v] + nd[v];
  for (vid s = 0; s < nv; ++s) {
    int64_t o = out_rowptr[s];
    for (int64_t k = rowptr[s]; k < rowptr[s + 1]; ++k) {
      vid d = col[k];
      if (deg[d] > deg[s] || (deg[d] == deg[s] && d > s)) out_col[o++] = d;
    }
  }
  return nnz;
}

// Set primitives exposed for tests (sequential semantics of the warp ops).
int32_t oracle_isect(const int32_t* a, int32_t na, const int32_t* b, int32_t nb, int32_t upper,
                     int32_t* out) {
  return isect(a, na, b, nb, out, upper);
}
uint64_t oracle_isect_num(const int32_t* a, int32_t na, const int32_t* b, int32_t nb, int32_t upper) {
  return isect_num(a, na, b, nb, upper);
}
// a − b restricted to values < upper (merge), the semantics of the CPU
// VertexSet::difference_buf without the owner-vertex drop
// (common/VertexSet.cc:21-89).  The GPU difference_* of the reference return
// 0 for an empty b (include/set_difference.cuh:44,99,166,242); this restates
// the set difference itself, a − ∅ = a, which the library implements.
int32_t oracle_diff(const int32_t* a, int32_t na, const int32_t* b, int32_t nb, int32_t upper, int32_t* out) {
  int32_t n = 0, i = 0, j = 0;
  while (i < na && a[i] < upper) {
    if (j >= nb || a[i] < b[j]) out[n++] = a[i++];
    else if (a[i] == b[j]) { ++i; ++j; }
    else ++j;
  }
  return n;
}

}  // extern "C"
