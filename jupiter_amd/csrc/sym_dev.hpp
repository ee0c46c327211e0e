// sym_dev.hpp — wave-level set operations and the per-edge contributions of
// P1..P12, shared by the single-query kernels and the batched / serving
// kernel.  See sym_kernels.hip for the formulas.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jk {
namespace symd {

constexpr int WPB = 4;        // waves per block, each an independent edge worker
constexpr int CAP = 256;      // LDS list capacity per wave (ints)
constexpr int STAGE = 256;    // longest search list staged into LDS
constexpr int NLIST = 5;      // per-wave lists: S, posS0, posS1, L, posL0
// block-per-edge kernel for the heavy edges of the q-sum patterns
constexpr int HV_WAVES = 16;           // waves per block (one edge)
constexpr int HV_NMAX = 8192;          // largest |S| it takes
constexpr int HV_WAVES_MID = 8;
constexpr int HV_BM_BITS = 1 << 18;    // graphs up to this many vertices: S as a bitmap
constexpr int HV_BM_HUGE = 1 << 20;    // ... and |S| > HV_NMAX runs in the huge shape
// |S| up to this: the narrow shape.  R-MAT 18, narrow width 512 / 1024 /
// 2048: P9 168 / 184 / 195 ms, P11 217 / 247 / 261, P5 438 / 407 / 424,
// P7 551 / 500 / 554, P12 511 / 456 / 494.
template <int P>
constexpr int heavy_mid() { return P == 9 || P == 11 ? 512 : 1024; }
// stream N(v2) when d(v2) <= HV_STREAM * |S|, else binary searches (R-MAT 18
// P9: ratio 2 / 4 / 8 / 16 / 32 / 64 -> 274 / 233 / 227 / 250 / 291 / 357 ms)
constexpr int HV_STREAM = 8;
template <int P>
constexpr bool heavy_split() { return P == 5 || P == 7 || P == 9 || P == 11 || P == 12; }

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ int64_t wsum64(int64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)((uint64_t)v >> 32);
    lo = __shfl_xor(lo, m);
    hi = __shfl_xor(hi, m);
    v += (int64_t)(((uint64_t)hi << 32) | lo);
  }
  return v;
}

__device__ __forceinline__ int wsum(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ int lb(const int32_t* a, int n, int32_t y) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < y) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// lb(a, n, y) computed by the whole wave (uniform arguments): each round
// splits the open range into at most 64 blocks and one load per lane finds
// the block, so a list of 25,000 takes 3 dependent loads instead of 15.
__device__ __forceinline__ int w_lb(const int32_t* a, int n, int32_t y) {
  const int ln = lane();
  int lo = 0, len = n;  // a[lo - 1] < y <= a[lo + len] (where they exist)
  while (len > 64) {
    const int step = (len + 63) >> 6;
    const int i = lo + ln * step;
    const int c = __popcll(__ballot(i < lo + len && a[i] < y));  // a prefix of the lanes
    if (c == 0) return lo;
    const int end = min(lo + c * step, lo + len);
    lo += (c - 1) * step + 1;
    len = end - lo;
  }
  return lo + __popcll(__ballot(ln < len && a[lo + ln] < y));
}

struct Wave {
  int32_t* stage;  // LDS, STAGE ints
};

// |A ∩ B| of two long sorted lists of comparable length (na <= nb <=
// MERGE_RATIO * na, na beyond the LDS staging size): a chunked merge.  Lane l
// takes the l-th of 64 contiguous chunks of A, finds where the chunk's first
// element falls in B by one binary search, then walks both lists forward as a
// sequential merge; every element of A is counted once, by the lane that owns
// it, and the B spans of neighbouring lanes meet at the chunk boundaries, so
// the wave reads about na + nb elements instead of na searches of log2(nb)
// dependent loads each (the merge side of the degree-ratio switch; the
// reference's CPU hybrid switches at ratio 32/50, common/intersect.cc:11,54).
constexpr int MERGE_RATIO = 8;
__device__ int64_t w_merge_count(const int32_t* A, int na, const int32_t* B, int nb) {
  const int chunk = (na + 63) >> 6;
  const int a0 = min(na, lane() * chunk), a1 = min(na, a0 + chunk);
  int c = 0;
  if (a0 < a1) {
    int j = lb(B, nb, A[a0]);
    for (int i = a0; i < a1 && j < nb; ++i) {
      const int32_t x = A[i];
      while (j < nb && B[j] < x) ++j;
      c += (j < nb && B[j] == x);
    }
  }
  return wsum(c);
}

// |A ∩ B| restricted to values < ub.  All lanes call with uniform arguments.
__device__ int64_t w_icount(Wave& w, const int32_t* A, int na, const int32_t* B, int nb, int32_t ub) {
  if (ub != INT32_MAX) {
    na = lb(A, na, ub);
    nb = lb(B, nb, ub);
  }
  if (na > nb) {
    const int32_t* t = A; A = B; B = t;
    int u = na; na = nb; nb = u;
  }
  if (na == 0) return 0;
  if (na > STAGE && nb <= MERGE_RATIO * na) return w_merge_count(A, na, B, nb);
  const int32_t* S = B;
  if (nb <= STAGE && na > 4) {
    for (int k = lane(); k < nb; k += 64) w.stage[k] = B[k];
    __syncwarp();
    S = w.stage;
  }
  int c = 0;
  for (int i = lane(); i < na; i += 64) {
    int32_t x = A[i];
    int p = lb(S, nb, x);
    c += (p < nb && S[p] == x);
  }
  c = wsum(c);
  __syncwarp();
  return c;
}

// out = A ∩ B (ascending, values < ub); posA/posB (optional) receive the
// element's index in A and B.  Returns the size.
__device__ int w_imat(Wave& w, const int32_t* A, int na, const int32_t* B, int nb, int32_t ub, int32_t* out,
                      int32_t* posA, int32_t* posB) {
  if (ub != INT32_MAX) {
    na = lb(A, na, ub);
    nb = lb(B, nb, ub);
  }
  const bool sw = na > nb;
  const int32_t* Lk = sw ? B : A;
  const int32_t* Sr = sw ? A : B;
  const int nl = sw ? nb : na;
  const int ns = sw ? na : nb;
  if (nl == 0) return 0;
  const int32_t* S = Sr;
  if (ns <= STAGE && nl > 4) {
    for (int k = lane(); k < ns; k += 64) w.stage[k] = Sr[k];
    __syncwarp();
    S = w.stage;
  }
  int n = 0;
  for (int base = 0; base < nl; base += 64) {
    const int i = base + lane();
    bool f = false;
    int32_t x = 0;
    int p = 0;
    if (i < nl) {
      x = Lk[i];
      p = lb(S, ns, x);
      f = p < ns && S[p] == x;
    }
    uint64_t m = __ballot(f);
    if (f) {
      int k = n + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      out[k] = x;
      if (posA) posA[k] = sw ? p : i;
      if (posB) posB[k] = sw ? i : p;
    }
    n += __popcll(m);
  }
  __syncwarp();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  return n;
}

// out = A − B restricted to values < ub: the elements of A (ascending) that B
// (ascending) does not hold, in order; out may be A itself (compaction in
// place).  A − ∅ = A (the reference's GPU difference_num / difference_set
// return 0 for an empty B, include/set_difference.cuh:44,99; no P-id reaches
// that case, and the vertex-induced plans need the set difference itself).
// Lanes take 64 elements of A at a time and binary-search B, staged in LDS
// when it is short; ballot + mbcnt keep the order.  Returns |out|.
__device__ int w_idiff(Wave& w, const int32_t* A, int na, const int32_t* B, int nb, int32_t ub, int32_t* out) {
  if (ub != INT32_MAX) {
    na = lb(A, na, ub);
    nb = lb(B, nb, ub);
  }
  const int32_t* S = B;
  if (nb > 0 && nb <= STAGE && na > 4) {
    for (int k = lane(); k < nb; k += 64) w.stage[k] = B[k];
    __syncwarp();
    S = w.stage;
  }
  int m = 0;
  for (int base = 0; base < na; base += 64) {
    const int i = base + lane();
    bool keep = false;
    int32_t x = 0;
    if (i < na) {
      x = A[i];
      const int p = nb > 0 ? lb(S, nb, x) : 0;
      keep = !(p < nb && S[p] == x);
    }
    const uint64_t b = __ballot(keep);
    if (keep)
      out[m + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = x;
    m += __popcll(b);
  }
  __syncwarp();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  return m;
}

__device__ __forceinline__ bool in_sorted(const int32_t* a, int n, int32_t x) {
  int p = lb(a, n, x);
  return p < n && a[p] == x;
}

struct Ctx {
  const int64_t* rp;
  const int32_t* ci;
  const int32_t* tri;
  const int64_t* nsum;
  __device__ const int32_t* N(int32_t v) const { return ci + rp[v]; }
  __device__ int d(int32_t v) const { return (int)(rp[v + 1] - rp[v]); }
};

__device__ __forceinline__ int64_t C2(int64_t x) { return x >= 2 ? x * (x - 1) / 2 : 0; }

// Contribution of directed edge (v0, v1) to pattern P (wave-uniform value).
template <int P>
__device__ int64_t edge_value(const Ctx& g, Wave& w, int32_t v0, int32_t v1, int32_t** L) {
  const int32_t* N0 = g.N(v0);
  const int32_t* N1 = g.N(v1);
  const int d0 = g.d(v0), d1 = g.d(v1);
  int32_t* S = L[0];
  int32_t* p0 = L[1];
  int32_t* p1 = L[2];
  int32_t* T = L[3];
  int32_t* pT = L[4];
  const int ln = lane();
  static_assert(P >= 2 && P <= 12, "P1 runs as wedge counting (p1_wedge_kernel)");
  if constexpr (P == 2) {
    if (v1 >= v0) return 0;
    return C2(w_icount(w, N0, d0, N1, d1, INT32_MAX));
  } else if constexpr (P == 3) {
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, p0, nullptr);
    int64_t sp = 0, sd = 0, st = 0;
    for (int k = ln; k < n; k += 64) {
      int32_t v2 = S[k];
      if (v2 > v1) {
        sp += 1;
        sd += g.d(v2) - 2;
        st += g.tri[g.rp[v0] + p0[k]] - 1;
      }
    }
    sp = wsum64(sp);
    sd = wsum64(sd);
    st = wsum64(st);
    return sp * (g.nsum[v0] - d1 - d0 + 1 - n) - sd - st;
  } else if constexpr (P == 4) {
    if (v1 >= v0) return 0;
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, p0, p1);
    if (n == 0) return 0;
    int64_t A = 0;
    for (int j = 0; j < d1; ++j) {
      int32_t v3 = N1[j];
      if (v3 == v0) continue;
      A += w_icount(w, N0, d0, g.N(v3), g.d(v3), INT32_MAX) - 1;
    }
    int64_t s0 = 0, s1 = 0;
    for (int k = ln; k < n; k += 64) {
      s0 += g.tri[g.rp[v0] + p0[k]] - 1;
      s1 += g.tri[g.rp[v1] + p1[k]] - 1;
    }
    return (int64_t)n * A - wsum64(s0) - wsum64(s1);
  } else if constexpr (P == 5) {
    if (v1 >= v0) return 0;
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, nullptr, nullptr);
    if (n < 3) return 0;
    int64_t k2 = 0;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      k2 += w_icount(w, S, n, g.N(v2), g.d(v2), INT32_MAX);
    }
    return (int64_t)(n - 2) * (k2 / 2);
  } else if constexpr (P == 6) {
    if (v1 >= v0) return 0;
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, nullptr, nullptr);
    int64_t c = 0;
    for (int i = 1; i < n; ++i) {
      int32_t v2 = S[i];
      for (int j = 0; j < i; ++j) {
        int32_t v3 = S[j];
        c += w_icount(w, g.N(v2), g.d(v2), g.N(v3), g.d(v3), INT32_MAX) - 2;
      }
    }
    return c;
  } else if constexpr (P == 7) {
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, p0, nullptr);
    int64_t c = 0;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      if (v2 <= v1) continue;
      int64_t q = w_icount(w, S, n, g.N(v2), g.d(v2), INT32_MAX);
      c += (int64_t)(n - 1) * (g.tri[g.rp[v0] + p0[j]] - 1) - q;
    }
    return c;
  } else if constexpr (P == 8) {
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, nullptr, nullptr);
    int64_t c = 0;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      if (v2 >= v1) break;  // S ascending
      const int m = w_imat(w, N0, d0, g.N(v2), g.d(v2), v1, T, nullptr, nullptr);
      for (int k = 0; k < m; ++k) {
        int32_t v3 = T[k];
        c += w_icount(w, S, n, g.N(v3), g.d(v3), v2);
      }
    }
    return c;
  } else if constexpr (P == 9) {
    if (v1 <= v0) return 0;
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, nullptr, nullptr);
    int64_t c = 0;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      if (v2 <= v1) continue;
      c += C2(w_icount(w, S, n, g.N(v2), g.d(v2), INT32_MAX));
    }
    return c;
  } else if constexpr (P == 10) {
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, nullptr, nullptr);
    int64_t c = 0;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      const int m = w_imat(w, N0, d0, g.N(v2), g.d(v2), INT32_MAX, T, pT, nullptr);
      for (int k = 0; k < m; ++k) {
        int32_t v3 = T[k];
        if (v3 <= v1) continue;
        const int64_t ins = in_sorted(S, n, v3) ? 1 : 0;
        const int64_t t03 = g.tri[g.rp[v0] + pT[k]];
        const int64_t q3 = w_icount(w, S, n, g.N(v3), g.d(v3), INT32_MAX);
        c += (n - 1 - ins) * (t03 - 1 - ins) - (q3 - 1);
      }
    }
    return c;
  } else if constexpr (P == 11) {
    if (v1 <= v0) return 0;
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, p0, p1);
    int64_t c = 0;
    const int64_t a = n - 1;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      if (v2 <= v1) continue;
      const int64_t q = w_icount(w, S, n, g.N(v2), g.d(v2), INT32_MAX);
      const int64_t b = g.tri[g.rp[v0] + p0[j]] - 1;
      const int64_t cc = g.tri[g.rp[v1] + p1[j]] - 1;
      c += a * b * cc - q * (a + b + cc) + 2 * q;
    }
    return c;
  } else if constexpr (P == 12) {
    if (v1 <= v0) return 0;
    const int n = w_imat(w, N0, d0, N1, d1, INT32_MAX, S, nullptr, nullptr);
    if (n < 3) return 0;
    int64_t c = 0;
    for (int j = 0; j < n; ++j) {
      int32_t v2 = S[j];
      c += C2(w_icount(w, S, n, g.N(v2), g.d(v2), INT32_MAX));
    }
    return (int64_t)(n - 3) * c;
  }
  return 0;
}

}  // namespace symd
}  // namespace jk
