// clique_kernels.hip — gfx950 kernels for triangle and k-clique counting
// (TC and P13..P16) on the degree-oriented DAG.
//
// Replaces the reference's P13_kernel..P16_kernel
// (src/message_passing_gpu_kernels.cuh:1208-1537) and the level-by-level
// pattern_task_generate / pattern_extend_pull_workloads path
// (src/gpu_kernels_codegen.cuh:6-95,261-410).  The count for root v0 equals
// the reference's sum over the DAG edges (v0, v1): the number of k-cliques
// whose DAG source is v0.
//
// Design (one root vertex v0 per task, L = N+(v0) sorted, d = |L|):
//  1. Build.  L is staged in LDS.  The out-lists N+(L[i]) of all i are
//     streamed as one flattened, coalesced sequence (a lane's element knows its
//     list by a binary search over the LDS segment offsets); each element is
//     looked up in L by binary search in LDS and, if present at position p,
//     bit p of row i is set (LDS / L2 atomic OR).  Row i is then the bitmap
//     of L ∩ N+(L[i]).  Lists much longer than d (ratio switch) are not
//     streamed: every element of L is binary-searched in them instead
//     (galloping side of the merge/gallop choice).
//  2. Count.  k-cliques with source v0 are v0 plus a (k-1)-clique of the
//     DAG on L given by the rows.  TC needs no rows (it counts the matches of
//     step 1).  With d <= 64 every row is one 64-bit word and lane i counts
//     the cliques starting at i with AND + popcount.  With d > 64, the rows
//     are multi-word; for each i the wave compacts S = row i, and either
//     counts with word-wise AND/popcount (k = 4) or builds (per lane) the compressed
//     |S| x |S| successor bitmap of the DAG inside S (k >= 5) and recurses
//     on single words.
// All arithmetic is integer; counts are exact uint64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace jk {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl_xor(lo, m);
  hi = __shfl_xor(hi, m);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
  return v;
}

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  return x;
}

// last index i in [0, n) with a[i] <= f, given a[0] <= f < a[n]
__device__ __forceinline__ int seg_search(const int* a, int n, int f) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= f) lo = mid; else hi = mid;
  }
  return lo;
}

// first index p in [0, n) with a[p] >= y (n if none)
template <class T>
__device__ __forceinline__ int lower_bound_dev(const T* a, int n, int y) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < y) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool bsearch_global(const int32_t* __restrict__ a, int64_t n, int32_t y) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    int32_t v = a[mid];
    if (v < y) lo = mid + 1; else hi = mid;
  }
  return lo < n && a[lo] == y;
}

// Group of G waves sharing one root (G == 1: a lone wave; G > 1: the block).
template <int G>
struct Grp {
  static constexpr int size = 64 * G;
  __device__ static int rank() { return G == 1 ? lane_id() : (int)threadIdx.x; }
  __device__ static int wave() { return G == 1 ? 0 : (int)(threadIdx.x >> 6); }
  __device__ static void sync() {
    if constexpr (G == 1) __syncwarp(); else __syncthreads();
  }
};

// Exclusive scan of a[0..n) in place (n <= 64 * G * per-thread chunk), sets a[n] = total.
template <int G>
__device__ void group_excl_scan(int* a, int n, int* wtot) {
  const int r = Grp<G>::rank();
  const int per = (n + Grp<G>::size - 1) / Grp<G>::size;
  const int b = r * per;
  const int e = min(n, b + per);
  int s = 0;
  for (int t = b; t < e; ++t) s += a[t];
  int incl = wave_incl_scan(s);
  int base = 0;
  if constexpr (G > 1) {
    const int w = Grp<G>::wave();
    if (lane_id() == 63) wtot[w] = incl;
    __syncthreads();
    for (int k = 0; k < w; ++k) base += wtot[k];
    if (w == G - 1 && lane_id() == 63) wtot[G] = base + incl;
  }
  int run = base + incl - s;
  for (int t = b; t < e; ++t) {
    int v = a[t];
    a[t] = run;
    run += v;
  }
  if constexpr (G == 1) {
    if (lane_id() == 63) a[n] = incl;
  }
  Grp<G>::sync();
  if constexpr (G > 1) {
    if (threadIdx.x == 0) a[n] = wtot[G];
    __syncthreads();
  }
}

// Count (depth)-cliques-with-prefix over single-word rows `R`: depth 1 is
// popcount(C); deeper levels branch on every member x of C with C & R[x].
// With BYTES the weights w[x] of every visited member are summed instead
// (work-model bytes of the reference nest).
template <int DEPTH, bool BYTES>
__device__ uint64_t rec1(uint64_t C, const uint64_t* R, const uint32_t* w) {
  if constexpr (DEPTH <= 0) {
    return 0;
  } else if constexpr (DEPTH == 1 && !BYTES) {
    return (uint64_t)__popcll(C);
  } else {
    uint64_t s = 0;
    if constexpr (!BYTES) {
      if (__popcll(C) < DEPTH) return 0;
    }
    // Positions follow vertex ids, not the DAG order, so a successor of x
    // may sit at a lower position: branch on C & R[x] with the whole of C.
    const uint64_t C0 = C;
    while (C) {
      int x = __builtin_ctzll(C);
      C &= C - 1;
      if constexpr (BYTES) s += w[x];
      if constexpr (DEPTH > 1) s += rec1<DEPTH - 1, BYTES>(C0 & R[x], R, w);
    }
    return s;
  }
}

// Multi-word fallback: words of the current set are the AND of the rows of
// the chosen positions p[0..NP), recomputed on the fly (no register arrays).
template <int DEPTH, int NP, bool BYTES, class RowT>
__device__ uint64_t recW(const RowT& R, int wlo, int whi, const int* p, const uint32_t* wt) {
  uint64_t s = 0;
  for (int w = wlo; w <= whi; ++w) {
    uint64_t m = R(p[0], w);
#pragma unroll
    for (int k = 1; k < NP; ++k) m &= R(p[k], w);
    if constexpr (DEPTH == 1 && !BYTES) {
      s += (uint64_t)__popcll(m);
    } else {
      while (m) {
        int x = (w << 6) | __builtin_ctzll(m);
        m &= m - 1;
        if constexpr (BYTES) s += wt[x];
        if constexpr (DEPTH > 1) {
          int q[NP + 1];
#pragma unroll
          for (int k = 0; k < NP; ++k) q[k] = p[k];
          q[NP] = x;
          s += recW<DEPTH - 1, NP + 1, BYTES>(R, wlo, whi, q, wt);
        }
      }
    }
  }
  return s;
}

struct RowsLDS {
  uint64_t* p;
  int W;
  __device__ uint64_t operator()(int i, int w) const { return p[i * W + w]; }
};
struct RowsGlobal {
  uint64_t* p;
  int W;
  __device__ uint64_t operator()(int i, int w) const { return __builtin_nontemporal_load(p + (int64_t)i * W + w); }
};

// Block reduction of a per-thread uint64 and one atomicAdd per block.
template <int NT>
__device__ void block_add(uint64_t v, uint64_t* out) {
  __shared__ uint64_t part[NT / 64];
  v = wave_sum_u64(v);
  if (lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) s += part[k];
    if (s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
  }
}

// ----------------------------------------------------------------------------
// Kernel.  G waves per root; WPB roots (groups) per block when G == 1.
// WMAX words per row => roots with d <= 64 * WMAX.  LROWS: rows in LDS,
// otherwise in a per-block slice of global scratch.
template <int G, int WPB, int WMAX, int K, bool BYTES, bool LROWS>
__global__ void __launch_bounds__(64 * G * WPB)
clique_kernel(CliqueArgs a) {
  constexpr int DM = 64 * WMAX;
  constexpr int NT = 64 * G * WPB;
  constexpr bool NEED_ROWS = (K >= 4);
  __shared__ int32_t s_L[WPB][DM];
  __shared__ int32_t s_off[WPB][DM + 1];
  __shared__ int64_t s_ub[WPB][DM];
  __shared__ int32_t s_hdu[WPB][DM];
  __shared__ uint32_t s_w[WPB][BYTES ? DM : 1];
  __shared__ uint64_t s_rows[WPB][(LROWS && NEED_ROWS) ? DM * WMAX : 1];
  __shared__ int32_t s_J[G * WPB][(WMAX > 1 && NEED_ROWS) ? DM : 1];
  __shared__ uint64_t s_c[G * WPB][(WMAX > 1 && K >= 5) ? 64 : 1];
  __shared__ uint32_t s_cw[G * WPB][(WMAX > 1 && K >= 5 && BYTES) ? 64 : 1];
  __shared__ int32_t s_wtot[G + 1];
  __shared__ int64_t s_task;

  const int slot = (G == 1) ? (int)(threadIdx.x >> 6) : 0;
  const int r = Grp<G>::rank();
  const int wv = Grp<G>::wave();
  const int ln = lane_id();
  int32_t* L = s_L[slot];
  int32_t* off = s_off[slot];
  int64_t* ub = s_ub[slot];
  int32_t* hdu = s_hdu[slot];
  uint32_t* wL = s_w[slot];
  uint64_t* rowp = LROWS ? s_rows[slot] : (a.gscr + (int64_t)blockIdx.x * DM * WMAX);
  const int64_t* __restrict__ rp = a.rp;
  const int32_t* __restrict__ ci = a.ci;

  uint64_t acc = 0;
  for (;;) {
    int64_t base;
    if constexpr (G == 1) {
      int64_t t = 0;
      if (ln == 0) t = (int64_t)atomicAdd(a.qctr, (unsigned)a.chunk);
      base = __shfl(t, 0);
    } else {
      if (threadIdx.x == 0) s_task = (int64_t)atomicAdd(a.qctr, 1u);
      __syncthreads();
      base = s_task;
      __syncthreads();
    }
    if (base >= a.ntasks) break;
    const int64_t tend = min(a.ntasks, base + (G == 1 ? (int64_t)a.chunk : 1));
    for (int64_t task = base; task < tend; ++task) {
      const int32_t v0 = a.ids ? a.ids[task] : (int32_t)(a.rb + task);
      const int64_t b0 = rp[v0];
      const int d = (int)(rp[v0 + 1] - b0);
      if (d < a.dmin || d > a.dmax || d > DM) continue;
      const int W = (d + 63) >> 6;
      if constexpr (BYTES) {
        if (r == 0) acc += 16 + 4 * (uint64_t)d;
      }
      // ---- stage root list and segment descriptors
      for (int t = r; t < d; t += Grp<G>::size) L[t] = ci[b0 + t];
      Grp<G>::sync();
      for (int t = r; t < d; t += Grp<G>::size) {
        int32_t u = L[t];
        int64_t s = rp[u];
        int du = (int)(rp[u + 1] - s);
        bool heavy = du > 64 && du > a.ratio * d;
        ub[t] = s;
        off[t] = heavy ? 0 : du;
        hdu[t] = heavy ? du : 0;
        if constexpr (BYTES) {
          wL[t] = 16u + 4u * (uint32_t)du;
          acc += 16 + 4 * (uint64_t)du;  // f(v1) of the edge (v0, L[t])
        }
      }
      if constexpr (NEED_ROWS) {
        for (int t = r; t < d * W; t += Grp<G>::size) rowp[t] = 0;
      }
      if constexpr (!LROWS && NEED_ROWS) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      Grp<G>::sync();
      if constexpr (K == 3 && BYTES) continue;  // TC work model needs no build
      group_excl_scan<G>(off, d, s_wtot);
      const int T = off[d];
      // ---- flattened stream of the light lists
      for (int f = r; f < T; f += Grp<G>::size) {
        int i = seg_search(off, d, f);
        int32_t y = ci[ub[i] + (f - off[i])];
        int p = lower_bound_dev(L, d, y);
        if (p < d && L[p] == y) {
          if constexpr (K == 3) {
            ++acc;
          } else {
            atomicOr((unsigned long long*)&rowp[i * W + (p >> 6)], 1ull << (p & 63));
          }
        }
      }
      // ---- long lists: binary search every element of L in them
      for (int i = 0; i < d; ++i) {
        const int du = hdu[i];
        if (du == 0) continue;
        const int32_t* lst = ci + ub[i];
        for (int t0 = wv * 64; t0 < d; t0 += Grp<G>::size) {
          int t = t0 + ln;
          bool f = t < d && bsearch_global(lst, du, L[t]);
          uint64_t m = __ballot(f);
          if constexpr (K == 3) {
            if (ln == 0) acc += (uint64_t)__popcll(m);
          } else {
            if (ln == 0 && m) atomicOr((unsigned long long*)&rowp[i * W + (t0 >> 6)], m);
          }
        }
      }
      if constexpr (!LROWS && NEED_ROWS) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      Grp<G>::sync();
      if constexpr (!LROWS && NEED_ROWS) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if constexpr (K == 3) continue;
      // ---- count
      if constexpr (WMAX == 1) {
        // one word per row, lane i owns row i
        if (r < d) {
          const uint64_t Ri = rowp[r];
          if constexpr (BYTES) acc += rec1<K - 3, true>(Ri, rowp, wL);
          else acc += rec1<K - 2, false>(Ri, rowp, wL);
        }
      } else {
        int32_t* J = s_J[G == 1 ? slot : wv];
        for (int i = wv; i < d; i += G) {
          // compact S = row i into J (ascending positions)
          int n = 0, wlo = W, whi = -1;
          for (int w = 0; w < W; ++w) {
            uint64_t word = LROWS ? rowp[i * W + w] : RowsGlobal{rowp, W}(i, w);
            if (!word) continue;
            wlo = min(wlo, w);
            whi = w;
            if ((word >> ln) & 1ull) {
              uint64_t below = ln ? (word & ((~0ull) >> (64 - ln))) : 0ull;
              J[n + __popcll(below)] = (w << 6) | ln;
            }
            n += __popcll(word);
          }
          __syncwarp();
          if (n == 0) continue;
          if constexpr (K == 4) {
            for (int t = ln; t < n; t += 64) {
              const int j = J[t];
              if constexpr (BYTES) {
                acc += wL[j];
              } else {
                uint64_t c = 0;
                for (int w = wlo; w <= whi; ++w) {
                  if constexpr (LROWS) c += __popcll(rowp[i * W + w] & rowp[j * W + w]);
                  else c += __popcll(RowsGlobal{rowp, W}(i, w) & RowsGlobal{rowp, W}(j, w));
                }
                acc += c;
              }
            }
          } else if (n <= 64) {
            // compressed |S| x |S| bitmap of the DAG inside S: lane t holds
            // row t = {s : J[s] in row J[t]} (successor sets, the reference's
            // extension direction), then recurses on single words.
            uint64_t* Rc = s_c[G == 1 ? slot : wv];
            uint32_t* wc = s_cw[G == 1 ? slot : wv];
            uint64_t mycol = 0;
            const int js = ln < n ? J[ln] : 0;
            int curw = -1;
            uint64_t word = 0;
            for (int t = 0; t < n; ++t) {
              const int x = J[t];
              if (ln < n) {
                if ((x >> 6) != curw) {
                  curw = x >> 6;
                  word = LROWS ? rowp[js * W + curw] : RowsGlobal{rowp, W}(js, curw);
                }
                mycol |= ((word >> (x & 63)) & 1ull) << t;
              }
            }
            if (ln < n) {
              Rc[ln] = mycol;
              if constexpr (BYTES) wc[ln] = wL[js];
            }
            __syncwarp();
            if (ln < n) {
              if constexpr (BYTES) acc += wc[ln] + rec1<K - 4, true>(mycol, Rc, wc);
              else acc += rec1<K - 3, false>(mycol, Rc, wc);
            }
            __syncwarp();
          } else {
            for (int t = ln; t < n; t += 64) {
              int p2[2] = {i, J[t]};
              if constexpr (BYTES) {
                acc += wL[J[t]];
                if constexpr (LROWS) acc += recW<K - 4, 2, true>(RowsLDS{rowp, W}, wlo, whi, p2, wL);
                else acc += recW<K - 4, 2, true>(RowsGlobal{rowp, W}, wlo, whi, p2, wL);
              } else {
                if constexpr (LROWS) acc += recW<K - 3, 2, false>(RowsLDS{rowp, W}, wlo, whi, p2, wL);
                else acc += recW<K - 3, 2, false>(RowsGlobal{rowp, W}, wlo, whi, p2, wL);
              }
            }
          }
          __syncwarp();
        }
      }
      Grp<G>::sync();
    }
  }
  block_add<NT>(acc, a.out);
}

// ---------------------------------------------------------------------------
// Host-side launch table.
template <int G, int WPB, int WMAX, int K, bool BYTES, bool LROWS>
static hipError_t launch_one(const CliqueArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL((clique_kernel<G, WPB, WMAX, K, BYTES, LROWS>), dim3(grid), dim3(64 * G * WPB), 0, st, a);
  return hipGetLastError();
}

template <int K, bool BYTES>
static hipError_t launch_k(int tier, const CliqueArgs& a, int grid, hipStream_t st) {
  switch (tier) {
    case TIER_WAVE:   return launch_one<1, 4, 1, K, BYTES, true>(a, grid, st);
    case TIER_LDS2:   return launch_one<4, 1, 2, K, BYTES, true>(a, grid, st);
    case TIER_LDS4:   return launch_one<4, 1, 4, K, BYTES, true>(a, grid, st);
    case TIER_LDS8:   return launch_one<4, 1, 8, K, BYTES, true>(a, grid, st);
    case TIER_GLB32:  return launch_one<4, 1, 32, K, BYTES, false>(a, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_clique(int k, bool bytes, int tier, const CliqueArgs& a, int grid, hipStream_t st) {
  switch (k) {
    case 3: return bytes ? launch_k<3, true>(tier, a, grid, st) : launch_k<3, false>(tier, a, grid, st);
    case 4: return bytes ? launch_k<4, true>(tier, a, grid, st) : launch_k<4, false>(tier, a, grid, st);
    case 5: return bytes ? launch_k<5, true>(tier, a, grid, st) : launch_k<5, false>(tier, a, grid, st);
    case 6: return bytes ? launch_k<6, true>(tier, a, grid, st) : launch_k<6, false>(tier, a, grid, st);
    case 7: return bytes ? launch_k<7, true>(tier, a, grid, st) : launch_k<7, false>(tier, a, grid, st);
  }
  return hipErrorInvalidValue;
}

int tier_max_deg(int tier) {
  switch (tier) {
    case TIER_WAVE: return 64;
    case TIER_LDS2: return 128;
    case TIER_LDS4: return 256;
    case TIER_LDS8: return 512;
    case TIER_GLB32: return 2048;
  }
  return 0;
}

size_t tier_scratch_words(int tier) {
  return tier == TIER_GLB32 ? (size_t)2048 * 32 : 0;
}

}  // namespace jk
