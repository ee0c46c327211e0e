// sym_kernels.hip — gfx950 kernels for the patterns P1..P12 on the symmetric
// CSR, plus the per-edge triangle counts and per-vertex sums they use.
//
// Replaces P1_kernel..P12_kernel (src/message_passing_gpu_kernels.cuh:104-1109)
// and the delegation producer/extend pairs of src/delegation_gpu_kernels.cuh.
// One wave owns one directed edge (v0, v1) of the root range, exactly the
// edge loop of the reference (eid over the local COO list, :22 etc.), and adds
// that edge's contribution to the count, so any root range gives the same
// partial count as the reference's per-GPU edge range.
//
// The contribution of an edge is the reference nest's value, but every
// innermost "for x in L: if x not excluded: count++" loop is replaced by an
// exact closed form in set sizes (|L| minus the excluded members that lie in
// L), and several nests are folded into sums over S = N(v0) ∩ N(v1) of
// per-edge triangle counts t(u,w) = |N(u) ∩ N(w)| and per-triangle common
// neighbour counts q(v2) = |S ∩ N(v2)|:
//   P2  C(s,2)                                   s = |S|         (v1 < v0)
//   P3  s'(D0 - d1 - d0 + 1 - s) - sum_{S'}(d(v2) - 2) - sum_{S'}(t(v0,v2) - 1)
//       with S' = {v2 in S : v2 > v1}, D0 = sum_{u in N(v0)} d(u)
//   P4  s*A - sum_S (t(v0,v3) - 1) - sum_S (t(v1,v2) - 1),
//       A = sum_{v3 in N(v1)\{v0}} (|N(v0) ∩ N(v3)| - 1)          (v1 < v0)
//   P5  (s - 2) * (sum_S q(v2)) / 2                               (v1 < v0)
//   P6  sum over pairs of S of (|N(v2) ∩ N(v3)| - 2)              (v1 < v0)
//   P7  sum_{S'} ((s - 1)(t(v0,v2) - 1) - q(v2))
//   P8  nest kept: v2 in S, v2 < v1; v3 in N(v0) ∩ N(v2), v3 < v1; |{x in S ∩ N(v3): x < v2}|
//   P9  sum_{S'} C(q(v2), 2)                                      (v0 < v1)
//   P10 v2 in S, v3 in N(v0) ∩ N(v2), v3 > v1:
//       (s - 1 - [v3 in S])(t(v0,v3) - 1 - [v3 in S]) - (|S ∩ N(v3)| - 1)
//   P11 sum_{S'} (a b c - q (a + b + c) + 2 q), a = s-1, b = t(v0,v2)-1, c = t(v1,v2)-1   (v0 < v1)
//   P12 (s - 3) sum_S C(q(v2), 2)                                 (v0 < v1)
// P1 runs as per-root wedge counting (p1_wedge_kernel below), which sums the
// reference nest's per-edge values exactly.
//
// Set operations are wave-wide: the shorter list is the lookup side, each
// lane binary-searches its element in the longer list (staged in LDS when it
// fits, the galloping/binary-search side of the merge/gallop choice),
// ballot + mbcnt compact the matches in order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clique_dev.hpp"  // seg_advance, wave_sum_u64
#include "kernels.hpp"
#include "sym_dev.hpp"

namespace jk {

using namespace symd;

// the orientation on which edge_value<P> can be non-zero
template <int P>
__device__ __forceinline__ bool heavy_orient(int32_t v0, int32_t v1) {
  if constexpr (P == 5) return v1 < v0;
  else if constexpr (P == 7) return true;
  else return v1 > v0;  // P9, P11, P12
}

template <int P>
__global__ void __launch_bounds__(64 * WPB) sym_kernel(SymArgs a) {
  __shared__ int32_t s_list[WPB][NLIST][CAP];
  __shared__ int32_t s_stage[WPB][STAGE];
  __shared__ uint64_t s_part[WPB];
  const int wv = threadIdx.x >> 6;
  const int ln = lane();
  Wave w{s_stage[wv]};
  Ctx g{a.rp, a.ci, a.tri, a.nsum};
  int32_t* glist = a.scr + ((int64_t)blockIdx.x * WPB + wv) * NLIST * a.scr_stride;
  uint64_t acc = 0;
  for (;;) {
    int64_t t = 0;
    if (ln == 0) t = (int64_t)atomicAdd((unsigned long long*)a.qctr, (unsigned long long)a.chunk);
    const int64_t base = __shfl(t, 0);
    if (base >= a.ne) break;
    const int64_t end = min(a.ne, base + (int64_t)a.chunk);
    for (int64_t e = base; e < end; ++e) {
      const int64_t eid = a.e0 + e;
      const int32_t v0 = a.src[eid];
      const int32_t v1 = a.ci[eid];
      if constexpr (heavy_split<P>()) {
        if (a.heavy_min > 0) {  // heavy: onto a list for sym_heavy_kernel
          const int32_t t = a.tri[eid];
          if (t > a.heavy_min && (t <= HV_NMAX || a.nv <= HV_BM_HUGE)) {
            if (!heavy_orient<P>(v0, v1)) continue;  // its value is 0
            const int wide = t > HV_NMAX ? 2 : t > a.heavy_mid;
            unsigned long long k = 0;
            if (ln == 0) k = atomicAdd(a.hctr + wide, 1ull);
            k = __shfl(k, 0);
            if (k < (unsigned long long)a.hcap) {
              if (ln == 0) a.hlist[(int64_t)wide * a.hcap + (int64_t)k] = (int32_t)(eid - a.e0);
              continue;
            }
          }
        }
      }
      // lists fit in LDS when the smaller endpoint degree does
      const int dmin = min((int)(a.rp[v0 + 1] - a.rp[v0]), (int)(a.rp[v1 + 1] - a.rp[v1]));
      int32_t* L[NLIST];
      const bool small = dmin <= CAP && (P != 8 && P != 10 ? true : (int)(a.rp[v0 + 1] - a.rp[v0]) <= CAP);
#pragma unroll
      for (int k = 0; k < NLIST; ++k) L[k] = small ? s_list[wv][k] : glist + (int64_t)k * a.scr_stride;
      int64_t v = edge_value<P>(g, w, v0, v1, L);
      if (ln == 0) acc += (uint64_t)v;
    }
  }
  if (ln == 0) s_part[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int k = 0; k < WPB; ++k) s += s_part[k];
    if (s) atomicAdd((unsigned long long*)a.out, (unsigned long long)s);
  }
}

// ---------------------------------------------------------------------------
// Whole-graph P3 over triangles.  The per-edge value above is a sum over
// (v0, v1, v2) with v0 v1 v2 a triangle and v2 > v1 of
//   X(v0, v1) - (d(v2) - 2) - (t(v0, v2) - 1),  X = D0 - d1 - d0 + 1 - t(v0, v1),
// and a triangle a < b < c meets v2 > v1 in exactly three orientations:
// (a, b, c), (b, a, c) and (c, a, b).  So one intersection per undirected
// edge (a, b), a < b, restricted to common neighbours c > b, finds every
// triangle once, and t(a, c), t(b, c) come from c's positions in N(a) and
// N(b) through the triangle array: half the intersections of the per-edge
// form, each over the members above b only.  Every root's share is in the
// total, so only a whole-graph query runs this way.
__global__ void __launch_bounds__(64 * WPB) p3_tri_kernel(SymArgs a) {
  __shared__ int32_t s_list[WPB][NLIST][CAP];
  __shared__ int32_t s_stage[WPB][STAGE];
  __shared__ int64_t s_part[WPB];
  const int wv = threadIdx.x >> 6;
  const int ln = lane();
  Wave w{s_stage[wv]};
  int32_t* glist = a.scr + ((int64_t)blockIdx.x * WPB + wv) * NLIST * a.scr_stride;
  int64_t acc = 0;  // the lane's terms
  for (;;) {
    int64_t t = 0;
    if (ln == 0) t = (int64_t)atomicAdd((unsigned long long*)a.qctr, (unsigned long long)a.chunk);
    const int64_t base = __shfl(t, 0);
    if (base >= a.ne) break;
    const int64_t end = min(a.ne, base + (int64_t)a.chunk);
    for (int64_t e = base; e < end; ++e) {
      const int64_t eid = a.e0 + e;
      const int32_t va = a.src[eid];
      const int32_t vb = a.ci[eid];
      if (vb <= va) continue;
      const int64_t ba = a.rp[va], bb = a.rp[vb];
      const int da = (int)(a.rp[va + 1] - ba), db = (int)(a.rp[vb + 1] - bb);
      const int32_t* Na = a.ci + ba;
      const int32_t* Nb = a.ci + bb;
      const int oa = w_lb(Na, da, vb + 1), ob = w_lb(Nb, db, vb + 1);  // members above b
      const int ma = da - oa, mb = db - ob;
      if (ma == 0 || mb == 0) continue;
      const bool small = min(ma, mb) <= CAP;
      int32_t* S = small ? s_list[wv][0] : glist;
      int32_t* pa = small ? s_list[wv][1] : glist + a.scr_stride;
      int32_t* pb = small ? s_list[wv][2] : glist + 2 * a.scr_stride;
      const int n = w_imat(w, Na + oa, ma, Nb + ob, mb, INT32_MAX, S, pa, pb);
      const int64_t tab = a.tri[eid];
      const int64_t xab = a.nsum[va] - db - da + 1 - tab;
      const int64_t xba = a.nsum[vb] - da - db + 1 - tab;
      for (int k = ln; k < n; k += 64) {
        const int32_t c = S[k];
        const int64_t tac = a.tri[ba + oa + pa[k]];
        const int64_t tbc = a.tri[bb + ob + pb[k]];
        const int64_t dc = a.rp[c + 1] - a.rp[c];
        acc += xab - (dc - 2) - (tac - 1);                                // (a, b, c)
        acc += xba - (dc - 2) - (tbc - 1);                                // (b, a, c)
        acc += (a.nsum[c] - da - dc + 1 - tac) - (db - 2) - (tbc - 1);   // (c, a, b)
      }
      __syncwarp();  // the lists are rewritten by the next edge
    }
  }
  acc = wsum64(acc);
  if (ln == 0) s_part[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int k = 0; k < WPB; ++k) s += s_part[k];
    if (s) atomicAdd((unsigned long long*)a.out, (unsigned long long)s);
  }
}

hipError_t launch_p3_triangles(const SymArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(p3_tri_kernel, dim3(grid), dim3(64 * WPB), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Heavy edges of P5, P7, P9, P11 and P12.  Their per-edge value sums terms
// of q(v2) = |S ∩ N(v2)| over the members v2 of S = N(v0) ∩ N(v1), and on
// skewed graphs nearly all the work sits in the few edges with a large S
// (R-MAT 18: the 12 % of directed edges with |S| > 128 hold 93 % of sum |S|^2).  A
// wave per edge computes each q(v2) as |S| binary searches in N(v2).  Here a
// block takes the edge: it intersects the two lists once, keeps S in LDS
// both as a list and as an open-addressing hash set, and its waves take the
// members v2 one at a time, counting q(v2) by streaming N(v2) (coalesced
// loads, one LDS probe sequence per element) when d(v2) <= HV_STREAM * |S|,
// else by binary searches of S's members in N(v2).  The per-edge value is the
// one edge_value<P> returns (the same formulas on the same q, b and c).

__device__ __forceinline__ uint32_t hv_hash(int32_t x, int lg) { return ((uint32_t)x * 2654435761u) >> (32 - lg); }

__device__ __forceinline__ int hv_has(const int32_t* tab, int mask, int lg, int32_t x) {
  uint32_t h = hv_hash(x, lg);
  for (;;) {
    const int32_t k = tab[h];
    if (k == x) return 1;
    if (k < 0) return 0;
    h = (h + 1) & (uint32_t)mask;
  }
}

// Three shapes, each popping the edges of its list one at a time: |S| <=
// NMAX of 512 or 1024 with 8 waves (small LDS, several blocks per CU), of
// 8192 with 16 waves (one block per CU), and (NMAX = 0) any |S| on graphs
// of at most HV_BM_HUGE vertices, S then in the block's slice of the list
// scratch and the set a bitmap of up to 128 KB.
template <int P, int NMAX, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) sym_heavy_kernel(SymArgs a) {
  constexpr bool HUGE_SHAPE = NMAX == 0;
  constexpr int TMAX = HUGE_SHAPE ? 0 : NMAX <= 1024 ? 4 * NMAX : 2 * NMAX;  // hash slots
  constexpr int BMW = (HUGE_SHAPE ? HV_BM_HUGE : HV_BM_BITS) / 32;
  // S as a set: a bitmap over the vertex ids when the graph has at most
  // HV_BM_BITS vertices (one LDS read per test), else the hash table
  __shared__ int32_t s_tab[TMAX > BMW ? TMAX : BMW];
  __shared__ int32_t s_Sl[HUGE_SHAPE ? 1 : NMAX];
  int32_t* const s_S = HUGE_SHAPE ? a.scr + (int64_t)blockIdx.x * a.scr_stride : s_Sl;
  __shared__ int64_t s_red[WAVES];
  __shared__ int64_t s_k;
  __shared__ int32_t s_n, s_next;
  constexpr int NT = 64 * WAVES;
  const int tid = threadIdx.x;
  const int wv = tid >> 6;
  const int ln = lane();
  const bool bm = HUGE_SHAPE || a.nv <= HV_BM_BITS;
  if (bm) {
    for (int k = tid; k < BMW; k += NT) s_tab[k] = 0;
  } else {
    for (int k = tid; k < TMAX; k += NT) s_tab[k] = -1;
  }
  // the list this shape takes: 0 (|S| <= heavy_mid), 1 or 2 (huge)
  const int wide = HUGE_SHAPE ? 2 : NMAX > a.heavy_mid;
  const int32_t* list = a.hlist + (int64_t)wide * a.hcap;
  const int64_t cnt = min((int64_t)a.hctr[wide], (int64_t)a.hcap);
  uint64_t acc = 0;  // thread 0's
  for (;;) {
    if (tid == 0) s_k = (int64_t)atomicAdd(a.hctr + 3 + wide, 1ull);
    __syncthreads();
    const int64_t k = s_k;
    if (k >= cnt) break;
    const int64_t eid = a.e0 + list[k];
    const int n = a.tri[eid];
    if (HUGE_SHAPE ? n <= a.scr_stride && a.nv <= HV_BM_HUGE : n <= NMAX) {  // always: listed by these bounds
      const int32_t v0 = a.src[eid], v1 = a.ci[eid];
      const int64_t b0 = a.rp[v0], b1 = a.rp[v1];
      const int d0 = (int)(a.rp[v0 + 1] - b0), d1 = (int)(a.rp[v1 + 1] - b1);
      const int32_t* N0 = a.ci + b0;
      const int32_t* N1 = a.ci + b1;
      // load <= 1/4 while the table allows (a miss, the common case, then
      // ends after ~1.4 probes), else <= 1/2
      int lg = 6;
      while ((1 << lg) < 4 * n && (1 << lg) < TMAX) ++lg;
      const int mask = (1 << lg) - 1;
      if (tid == 0) {
        s_n = 0;
        s_next = 0;
      }
      __syncthreads();
      // S: the shorter list's elements searched in the longer one
      const int32_t* A = d0 <= d1 ? N0 : N1;
      const int32_t* B = d0 <= d1 ? N1 : N0;
      const int na = min(d0, d1), nb = max(d0, d1);
      for (int i0 = 0; i0 < na; i0 += NT) {
        const int i = i0 + tid;
        int32_t x = 0;
        bool f = false;
        if (i < na) {
          x = A[i];
          const int p = lb(B, nb, x);
          f = p < nb && B[p] == x;
        }
        const uint64_t mk = __ballot(f);
        int kb = 0;
        if (ln == 0 && mk) kb = atomicAdd(&s_n, (int)__popcll(mk));
        kb = __shfl(kb, 0);
        const int k = kb + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
        if (f && k < n) {  // k < n always: the triangle array holds |S|
          s_S[k] = x;
          if (bm) {
            atomicOr(reinterpret_cast<uint32_t*>(&s_tab[x >> 5]), 1u << (x & 31));
          } else {
            uint32_t h = hv_hash(x, lg);
            while (atomicCAS(&s_tab[h], -1, x) != -1) h = (h + 1) & (uint32_t)mask;
          }
        }
      }
      __syncthreads();
      const int nS = min(n, s_n);  // = n: the members written to s_S
      // q(v2) for the members, one wave per member at a time
      int64_t part = 0;  // lane 0's
      for (;;) {
        int j = 0;
        if (ln == 0) j = atomicAdd(&s_next, 1);
        j = __shfl(j, 0);
        if (j >= nS) break;
        const int32_t v2 = s_S[j];
        if constexpr (P == 7 || P == 9 || P == 11) {
          if (v2 <= v1) continue;
        }
        int64_t r2 = a.rp[v2];
        int d2 = (int)(a.rp[v2 + 1] - r2);
        if constexpr (P == 5) {
          // P5 needs only sum_S q = 2 |E(S)|: each edge of S is taken once,
          // from its smaller end, so only N(v2)'s members above v2
          const int lo = w_lb(a.ci + r2, d2, v2 + 1);
          r2 += lo;
          d2 -= lo;
        }
        const int32_t* N2 = a.ci + r2;
        int q = 0;
        if (d2 <= a.heavy_stream * n) {
          int k = ln;
          if (bm) {
            const uint32_t* w = reinterpret_cast<const uint32_t*>(s_tab);
            for (; k + 192 < d2; k += 256) {  // four loads in flight
              const int32_t x0 = N2[k], x1 = N2[k + 64], x2 = N2[k + 128], x3 = N2[k + 192];
              q += ((w[x0 >> 5] >> (x0 & 31)) & 1) + ((w[x1 >> 5] >> (x1 & 31)) & 1) +
                   ((w[x2 >> 5] >> (x2 & 31)) & 1) + ((w[x3 >> 5] >> (x3 & 31)) & 1);
            }
            for (; k < d2; k += 64) {
              const int32_t x = N2[k];
              q += (w[x >> 5] >> (x & 31)) & 1;
            }
          } else {
            for (; k + 192 < d2; k += 256) {
              const int32_t x0 = N2[k], x1 = N2[k + 64], x2 = N2[k + 128], x3 = N2[k + 192];
              q += hv_has(s_tab, mask, lg, x0) + hv_has(s_tab, mask, lg, x1) + hv_has(s_tab, mask, lg, x2) +
                   hv_has(s_tab, mask, lg, x3);
            }
            for (; k < d2; k += 64) q += hv_has(s_tab, mask, lg, N2[k]);
          }
        } else {
          for (int k = ln; k < nS; k += 64) q += in_sorted(N2, d2, s_S[k]);  // (P5: members > v2 only)
        }
        q = wsum(q);
        // t(v0, v2) and t(v1, v2) from v2's positions in N(v0) and N(v1)
        int64_t bb = 0, cc = 0;
        if constexpr (P == 7 || P == 11) bb = a.tri[b0 + w_lb(N0, d0, v2)] - 1;
        if constexpr (P == 11) cc = a.tri[b1 + w_lb(N1, d1, v2)] - 1;
        if (ln == 0) {
          if constexpr (P == 5) {
            part += q;
          } else if constexpr (P == 9 || P == 12) {
            part += C2(q);
          } else if constexpr (P == 7) {
            part += (int64_t)(n - 1) * bb - q;
          } else {  // P11
            const int64_t aa = n - 1;
            part += aa * bb * cc - (int64_t)q * (aa + bb + cc) + 2 * (int64_t)q;
          }
        }
      }
      if (ln == 0) s_red[wv] = part;
      __syncthreads();
      if (tid == 0) {
        int64_t c = 0;
        for (int k = 0; k < WAVES; ++k) c += s_red[k];
        if constexpr (P == 5) c = n >= 3 ? (int64_t)(n - 2) * c : 0;  // c = |E(S)|
        else if constexpr (P == 12) c = n >= 3 ? (int64_t)(n - 3) * c : 0;
        acc += (uint64_t)c;
      }
      if (bm) {
        for (int k = tid; k < nS; k += NT) s_tab[s_S[k] >> 5] = 0;
      } else {
        for (int k = tid; k <= mask; k += NT) s_tab[k] = -1;
      }
      __syncthreads();
    } else {
      __syncthreads();  // s_k read by all before it is taken again
    }
  }
  if (tid == 0 && acc) atomicAdd((unsigned long long*)a.out, (unsigned long long)acc);
}

// ---------------------------------------------------------------------------
// P1 (4-cycles) by wedges.  The reference's per-edge value (edges (v0, v1),
// v1 < v0: sum over v2 in N(v0), v2 < v1, of |N(v1) ∩ N(v2) ∩ [0, v0)|;
// src/message_passing_gpu_kernels.cuh:104-186) summed over a root's edges is
//   sum over x < v0 of C(c_x, 2),  c_x = |{v in N(v0) : v < v0, x in N(v)}|:
// every pair v2 < v1 of the c_x such v with the common neighbour x is one
// term.  So a root needs one pass over its wedges (v0, v, x) with v, x < v0,
// not an intersection per pair.  One wave per root; the wedges of 64
// neighbours v at a time form one flattened stream (consecutive lanes read
// consecutive elements).  The root's wedge count W picks where the c_x live:
//  * W <= P1_HS / 2: the wave's LDS hash table, its first 2^t slots with
//    2^t >= 2W: open addressing on 64-bit slots (key x in the high half,
//    count in the low half; a slot is claimed by compare-and-swap and counted
//    by a 64-bit add), summed and emptied in one sweep of those 2^t slots;
//  * larger W up to a bound: a hash table of the same form over the first
//    words of the wave's global counter region (keys x + 1, so that zero is
//    the empty slot), claimed and counted by atomics, emptied by atomicExch;
//  * otherwise the region as dense counters: a first pass adds 1 to c_x, a
//    second takes every counter back to zero with atomicExch and adds C(c, 2)
//    for the value it took (a repeated x takes 0 after its first exchange).
constexpr uint64_t P1_EMPTY = 0xffffffff00000000ull;  // key -1, count 0

// Calls f(x) for every wedge endpoint x of root v0 (its m0 neighbours below
// v0 start at ci[b0]), 64 neighbours at a time.
template <class F>
__device__ void p1_wedges(const int64_t* rp, const int32_t* ci, int64_t b0, int m0, int32_t v0, int32_t* off,
                          int64_t* lbase, F f, int len0 = -1, int64_t base0 = 0) {
  const int ln = lane();
  for (int jb = 0; jb < m0; jb += 64) {
    const int j = jb + ln;
    int len = 0;
    int64_t base = 0;
    if (jb == 0 && len0 >= 0) {
      // the first 64 neighbours' segments, found by the caller
      len = len0;
      base = base0;
    } else if (j < m0) {
      const int32_t v = ci[b0 + j];
      base = rp[v];
      len = lb(ci + base, (int)(rp[v + 1] - base), v0);  // x < v0
    }
    int incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (ln >= o) incl += y;
    }
    const int T = __shfl(incl, 63);
    off[ln] = incl - len;
    lbase[ln] = base - (incl - len);  // element f of v's segment is ci[lbase + f]
    if (ln == 63) off[64] = T;
    __syncwarp();
    int i = 0;
    for (int fe = ln; fe < T; fe += 64) {
      i = seg_advance(off, 64, i, fe);
      f(ci[lbase[i] + fe]);
    }
    __syncwarp();
  }
}

// The 4-cycle count of root v0: sum over x < v0 of C(c_x, 2), the c_x in the
// wave's LDS table `tab` (roots with at most P1_HS / 2 wedges), else in a hash
// table laid over the wave's global counter region `cnt` (cnt_len ints, zero
// before, left zero), or in the region as dense counters when the table does
// not fit it.  Wave-uniform result.
template <int LOG>
__device__ uint64_t p1_root(const int64_t* rp, const int32_t* ci, int32_t v0, uint64_t* tab, int32_t* off,
                            int64_t* lbase, int32_t* cnt, int64_t cnt_len) {
  constexpr int P1_HS = 1 << LOG;
  constexpr int P1_LOG_HS = LOG;
  const int ln = lane();
  const int64_t b0 = rp[v0];
  const int m0 = lb(ci + b0, (int)(rp[v0 + 1] - b0), v0);  // neighbours below v0
  // wedge count of the root, to pick the table; the first 64 neighbours'
  // segments are kept for the wedge passes (most roots have no more), so
  // their binary searches are not repeated
  int64_t w = 0;
  int len0 = 0;
  int64_t base0 = 0;
  for (int j = ln; j < m0; j += 64) {
    const int32_t v = ci[b0 + j];
    const int64_t base = rp[v];
    const int len = lb(ci + base, (int)(rp[v + 1] - base), v0);
    if (j == ln) {
      len0 = len;
      base0 = base;
    }
    w += len;
  }
  w = wsum64(w);
  uint64_t acc = 0;
  if (w < 2) return 0;  // no pair of wedges shares an endpoint
  if (w <= P1_HS / 2) {
    // only the first 2^lt slots (load <= 1/2) are used, so the sweep below
    // costs the root's own wedge count, not the whole table
    int lt = 6;
    while (lt < P1_LOG_HS && ((int64_t)1 << lt) < 2 * w) ++lt;
    const int used = 1 << lt;
    p1_wedges(rp, ci, b0, m0, v0, off, lbase, [&](int32_t x) {
      const uint64_t key = (uint64_t)(uint32_t)x << 32;
      int h = (int)(((uint32_t)x * 2654435761u) >> (32 - lt));
      for (;;) {
        uint64_t cur = tab[h];
        if ((cur & P1_EMPTY) == key) {
          atomicAdd((unsigned long long*)&tab[h], 1ull);
          break;
        }
        if (cur == P1_EMPTY) {
          cur = atomicCAS((unsigned long long*)&tab[h], P1_EMPTY, key | 1ull);
          if (cur == P1_EMPTY) break;
          if ((cur & P1_EMPTY) == key) {  // claimed by another lane for the same x
            atomicAdd((unsigned long long*)&tab[h], 1ull);
            break;
          }
        }
        h = (h + 1) & (used - 1);
      }
    }, len0, base0);
    __syncwarp();
    for (int k = ln; k < used; k += 64) {
      const uint64_t v = tab[k];
      if (v != P1_EMPTY) {
        const uint64_t c = (uint32_t)v;
        acc += c * (c - 1) / 2;
        tab[k] = P1_EMPTY;
      }
    }
    __syncwarp();
  } else {
    // a hash table of 2^lt 64-bit slots (load <= 1/2) laid over the wave's
    // counter region when it is small and fits: key x + 1 in the high half (0 = empty, so
    // the region's all-zero state is an empty table), count in the low half.
    // Its slots are a compact block of the region, where the dense counters
    // below scatter over all nv of them.  Every access is an atomic, so no
    // stale cached line is read, and the sweep leaves the region zero.
    int lt = P1_LOG_HS + 1;
    while (((int64_t)1 << lt) < 2 * w) ++lt;
    const int64_t used = (int64_t)1 << lt;
    const uintptr_t a0 = (reinterpret_cast<uintptr_t>(cnt) + 7) & ~(uintptr_t)7;
    uint64_t* gt = reinterpret_cast<uint64_t*>(a0);
    const int64_t slots = ((reinterpret_cast<uintptr_t>(cnt + cnt_len) - a0) / 8);
    // beyond 2^15 slots (256 KiB) the table no longer stays in cache and its
    // probes cost more than the dense counters' (R-MAT 20, vertex order: half
    // the roots 1.02 s with dense counters, 1.50 s with unbounded tables)
    if (lt <= 15 && used <= slots) {
      p1_wedges(rp, ci, b0, m0, v0, off, lbase, [&](int32_t x) {
        const uint64_t key = (uint64_t)(uint32_t)(x + 1) << 32;
        uint64_t h = ((uint64_t)((uint32_t)x * 2654435761u) << lt) >> 32;  // top lt bits of the product
        for (;;) {
          const uint64_t cur = atomicCAS((unsigned long long*)&gt[h], 0ull, (unsigned long long)(key | 1ull));
          if (cur == 0ull) break;
          if ((cur >> 32) == (key >> 32)) {
            atomicAdd((unsigned long long*)&gt[h], 1ull);
            break;
          }
          h = (h + 1) & (uint64_t)(used - 1);
        }
      }, len0, base0);
      __threadfence();
      for (int64_t k = ln; k < used; k += 64) {
        const uint64_t v = atomicExch((unsigned long long*)&gt[k], 0ull);
        const uint64_t c = (uint32_t)v;
        acc += c * (c - 1) / 2;
      }
      __threadfence();
    } else {
      p1_wedges(rp, ci, b0, m0, v0, off, lbase, [&](int32_t x) { atomicAdd(&cnt[x], 1); }, len0, base0);
      __threadfence();  // the increments are performed before they are taken back
      p1_wedges(rp, ci, b0, m0, v0, off, lbase, [&](int32_t x) {
        const uint64_t c = (uint32_t)atomicExch(&cnt[x], 0);
        acc += c * (c - 1) / 2;
      }, len0, base0);
      __threadfence();
    }
  }
  return wave_sum_u64(acc);
}

template <int LOG>
__global__ void __launch_bounds__(64 * P1_WPB) p1_wedge_kernel(P1Args a) {
  constexpr int P1_HS = 1 << LOG;
  __shared__ uint64_t s_tab[P1_WPB][P1_HS];
  __shared__ int32_t s_off[P1_WPB][65];
  __shared__ int64_t s_lb[P1_WPB][64];
  const int wv = threadIdx.x >> 6;
  const int ln = lane();
  uint64_t* tab = s_tab[wv];
  int32_t* cnt = a.cnt + ((int64_t)blockIdx.x * P1_WPB + wv) * a.nv;
  for (int k = ln; k < P1_HS; k += 64) tab[k] = P1_EMPTY;
  __syncwarp();
  uint64_t acc = 0;
  for (;;) {
    int64_t t = 0;
    if (ln == 0) t = (int64_t)atomicAdd(a.qctr, 1u);
    t = __shfl(t, 0);
    if (a.rb + t >= a.re) break;
    const int64_t r = a.desc ? a.re - 1 - t : a.rb + t;
    acc += p1_root<LOG>(a.rp, a.ci, (int32_t)r, tab, s_off[wv], s_lb[wv], cnt, a.nv);
  }
  if (ln == 0 && acc) atomicAdd((unsigned long long*)a.out, (unsigned long long)acc);
}

// The 4-cycle queries of a batch in one launch: the roots of every query
// with kind 101 form one flattened task range (prefix p1cs over the query
// table); a wave takes one root at a time, finds its query by binary search
// over the prefix, and adds the root's count to the query's slot.
template <int LOG>
__global__ void __launch_bounds__(64 * P1_WPB) p1_batch_kernel(P1BatchArgs a) {
  constexpr int P1_HS = 1 << LOG;
  __shared__ uint64_t s_tab[P1_WPB][P1_HS];
  __shared__ int32_t s_off[P1_WPB][65];
  __shared__ int64_t s_lb[P1_WPB][64];
  const int wv = threadIdx.x >> 6;
  const int ln = lane();
  uint64_t* tab = s_tab[wv];
  int32_t* cnt = a.cnt + ((int64_t)blockIdx.x * P1_WPB + wv) * a.cnt_stride;
  for (int k = ln; k < P1_HS; k += 64) tab[k] = P1_EMPTY;
  __syncwarp();
  for (;;) {
    int64_t t = 0;
    if (ln == 0) t = (int64_t)atomicAdd(a.qctr, 1u);
    t = __shfl(t, 0);
    if (t >= a.total) break;
    int lo = 0, hi = a.nq;  // last query i with p1cs[i] <= t
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.p1cs[mid] <= t) lo = mid; else hi = mid;
    }
    // (ascending within a query: dealing a relabelled query's hubs first
    // measured slower here, 276 -> 284 ms for 256 queries, the batch's other
    // queries already filling the launch's tail)
    const BQ& Q = a.q[lo];
    const uint64_t c =
        p1_root<LOG>(Q.rp, Q.ci, (int32_t)(Q.base + (t - a.p1cs[lo])), tab, s_off[wv], s_lb[wv], cnt, a.cnt_stride);
    if (ln == 0 && c) atomicAdd((unsigned long long*)(a.counts + lo), (unsigned long long)c);
  }
}

// log2 of the per-wave LDS table slots: 4096 (32 KiB per wave) measured
// fastest against 2048 and 1024 on a batch of 256 Erdos-Renyi / Barabasi-
// Albert 4-cycle queries (323 vs 420 vs 516 ms), though the smaller tables
// give more waves per CU (R-MAT 18/20 and the 3.8M Chung-Lu graph: within
// 0-28%, the Chung-Lu graph faster with 1024).
constexpr int P1_TABLE_LOG2 = 12;

hipError_t launch_p1_batch(const P1BatchArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(p1_batch_kernel<P1_TABLE_LOG2>, dim3(grid), dim3(64 * P1_WPB), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_p1_wedge(const P1Args& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(p1_wedge_kernel<P1_TABLE_LOG2>, dim3(grid), dim3(64 * P1_WPB), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Auxiliary per-graph arrays.

// src[e] = source vertex of directed edge e (the reference's COO src list,
// src/jupiter_gpm.cu:129-146); one wave per vertex.
__global__ void k_fill_src(const int64_t* rp, int32_t nv, int32_t* src) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wid; v < nv; v += nw)
    for (int64_t e = rp[v] + lane(); e < rp[v + 1]; e += 64) src[e] = (int32_t)v;
}

// tri[e] = |N(u) ∩ N(v)| for every directed edge e = (u, v): one wave per edge.
__global__ void __launch_bounds__(256) k_edge_tri(const int64_t* rp, const int32_t* ci, const int32_t* src,
                                                  int64_t ne, int32_t* tri, unsigned long long* qctr) {
  __shared__ int32_t s_stage[4][STAGE];
  Wave w{s_stage[threadIdx.x >> 6]};
  const int ln = lane();
  for (;;) {
    int64_t t = 0;
    if (ln == 0) t = (int64_t)atomicAdd(qctr, 64ull);
    const int64_t base = __shfl(t, 0);
    if (base >= ne) break;
    const int64_t end = min(ne, base + 64);
    for (int64_t e = base; e < end; ++e) {
      // |N(u) ∩ N(v)| is the same for both directions: intersect once, from
      // the smaller end, and write it to (v, u) too (found by a wave-wide
      // lower bound in N(v))
      const int32_t u = src[e], v = ci[e];
      if (v < u) continue;
      const int64_t bv = rp[v];
      const int dv = (int)(rp[v + 1] - bv);
      int64_t c = w_icount(w, ci + rp[u], (int)(rp[u + 1] - rp[u]), ci + bv, dv, INT32_MAX);
      const int pr = w_lb(ci + bv, dv, u);
      if (ln == 0) {
        tri[e] = (int32_t)c;
        tri[bv + pr] = (int32_t)c;
      }
    }
  }
}

// nsum[v] = sum of the degrees of v's neighbours.
__global__ void k_nsum(const int64_t* rp, const int32_t* ci, int32_t nv, int64_t* nsum) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wid; v < nv; v += nw) {
    int64_t s = 0;
    for (int64_t e = rp[v] + lane(); e < rp[v + 1]; e += 64) {
      int32_t u = ci[e];
      s += rp[u + 1] - rp[u];
    }
    s = wsum64(s);
    if (lane() == 0) nsum[v] = s;
  }
}

// Per-vertex pattern sums over the root range: 3-stars C(d,3) and tailed
// triangles t_v (d_v - 2) with t_v = (sum of tri over v's edges) / 2.
__global__ void k_vertex_sums(const int64_t* rp, const int32_t* tri, int64_t rb, int64_t re, int mode,
                              unsigned long long* out) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t acc = 0;
  for (int64_t v = rb + wid; v < re; v += nw) {
    const int64_t d = rp[v + 1] - rp[v];
    if (mode == 0) {
      if (lane() == 0 && d >= 3) acc += (uint64_t)(d * (d - 1) / 2) * (uint64_t)(d - 2) / 3;
    } else {
      int64_t s = 0;
      for (int64_t e = rp[v] + lane(); e < rp[v + 1]; e += 64) s += tri[e];
      s = wsum64(s);
      if (lane() == 0 && d >= 2) acc += (uint64_t)(s / 2) * (uint64_t)(d - 2);
    }
  }
  acc = (uint64_t)wsum64((int64_t)acc);
  if (lane() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

// Diamonds (P2) over the directed edges [e0, e1) of a root range, from the
// per-edge triangle array: sum of C(tri[e], 2) over the edges (v0, v1) with
// v1 < v0, v0 = src[e].  One pass over three int32 streams, four edges per
// thread per step (16-byte loads where the range allows), so the pass runs at
// memory speed.
__global__ void __launch_bounds__(256) k_p2_edges(const int32_t* __restrict__ src, const int32_t* __restrict__ ci,
                                                  const int32_t* __restrict__ tri, int64_t e0, int64_t e1,
                                                  unsigned long long* out) {
  uint64_t acc = 0;
  const int64_t nt = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // head up to the first multiple of 4, then aligned quads, then the tail
  const int64_t a0 = min(e1, (e0 + 3) & ~(int64_t)3);
  const int64_t a1 = max(a0, e1 & ~(int64_t)3);
  for (int64_t e = e0 + tid; e < a0; e += nt)
    if (ci[e] < src[e]) acc += (uint64_t)C2(tri[e]);
  for (int64_t q = a0 / 4 + tid; q < a1 / 4; q += nt) {
    const int4 s4 = reinterpret_cast<const int4*>(src)[q];
    const int4 c4 = reinterpret_cast<const int4*>(ci)[q];
    const int4 t4 = reinterpret_cast<const int4*>(tri)[q];
    if (c4.x < s4.x) acc += (uint64_t)C2(t4.x);
    if (c4.y < s4.y) acc += (uint64_t)C2(t4.y);
    if (c4.z < s4.z) acc += (uint64_t)C2(t4.z);
    if (c4.w < s4.w) acc += (uint64_t)C2(t4.w);
  }
  for (int64_t e = a1 + tid; e < e1; e += nt)
    if (ci[e] < src[e]) acc += (uint64_t)C2(tri[e]);
  acc = (uint64_t)wsum64((int64_t)acc);
  __shared__ uint64_t part[4];
  if (lane() == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t s = part[0] + part[1] + part[2] + part[3];
    if (s) atomicAdd(out, (unsigned long long)s);
  }
}

hipError_t launch_p2_edges(const int32_t* src, const int32_t* ci, const int32_t* tri, int64_t e0, int64_t e1,
                           unsigned long long* out, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_p2_edges, dim3(grid), dim3(256), 0, st, src, ci, tri, e0, e1, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
template <int P>
static hipError_t launch_p(const SymArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(sym_kernel<P>, dim3(grid), dim3(64 * WPB), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_sym(int p, const SymArgs& a, int grid, hipStream_t st) {
  switch (p) {
    case 2: return launch_p<2>(a, grid, st);
    case 3: return launch_p<3>(a, grid, st);
    case 4: return launch_p<4>(a, grid, st);
    case 5: return launch_p<5>(a, grid, st);
    case 6: return launch_p<6>(a, grid, st);
    case 7: return launch_p<7>(a, grid, st);
    case 8: return launch_p<8>(a, grid, st);
    case 9: return launch_p<9>(a, grid, st);
    case 10: return launch_p<10>(a, grid, st);
    case 11: return launch_p<11>(a, grid, st);
    case 12: return launch_p<12>(a, grid, st);
  }
  return hipErrorInvalidValue;
}

template <int P>
static hipError_t launch_heavy_p(const SymArgs& a, int shape, int grid, hipStream_t st) {
  if (shape == 2) hipLaunchKernelGGL((sym_heavy_kernel<P, 0, HV_WAVES>), dim3(grid), dim3(64 * HV_WAVES), 0, st, a);
  else if (shape == 1) hipLaunchKernelGGL((sym_heavy_kernel<P, HV_NMAX, HV_WAVES>), dim3(grid), dim3(64 * HV_WAVES), 0, st, a);
  else hipLaunchKernelGGL((sym_heavy_kernel<P, heavy_mid<P>(), HV_WAVES_MID>), dim3(grid), dim3(64 * HV_WAVES_MID), 0, st, a);
  return hipGetLastError();
}

bool sym_heavy_pattern(int p) { return p == 5 || p == 7 || p == 9 || p == 11 || p == 12; }
int sym_heavy_nmax() { return HV_NMAX; }
int sym_heavy_stream() { return HV_STREAM; }

int sym_heavy_mid(int p) { return p == 9 || p == 11 ? heavy_mid<9>() : heavy_mid<5>(); }

hipError_t launch_sym_heavy(int p, const SymArgs& a, int shape, int grid, hipStream_t st) {
  switch (p) {
    case 5: return launch_heavy_p<5>(a, shape, grid, st);
    case 7: return launch_heavy_p<7>(a, shape, grid, st);
    case 9: return launch_heavy_p<9>(a, shape, grid, st);
    case 11: return launch_heavy_p<11>(a, shape, grid, st);
    case 12: return launch_heavy_p<12>(a, shape, grid, st);
  }
  return hipErrorInvalidValue;
}

int sym_waves_per_block() { return WPB; }
int sym_lists_per_wave() { return NLIST; }

hipError_t launch_aux_src(const int64_t* rp, int32_t nv, int32_t* src, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_src, dim3(grid), dim3(256), 0, st, rp, nv, src);
  return hipGetLastError();
}
hipError_t launch_aux_tri(const int64_t* rp, const int32_t* ci, const int32_t* src, int64_t ne, int32_t* tri,
                          unsigned long long* qctr, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_edge_tri, dim3(grid), dim3(256), 0, st, rp, ci, src, ne, tri, qctr);
  return hipGetLastError();
}
hipError_t launch_aux_nsum(const int64_t* rp, const int32_t* ci, int32_t nv, int64_t* nsum, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_nsum, dim3(grid), dim3(256), 0, st, rp, ci, nv, nsum);
  return hipGetLastError();
}
hipError_t launch_vertex_sums(const int64_t* rp, const int32_t* tri, int64_t rb, int64_t re, int mode,
                              unsigned long long* out, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_vertex_sums, dim3(grid), dim3(256), 0, st, rp, tri, rb, re, mode, out);
  return hipGetLastError();
}

}  // namespace jk
