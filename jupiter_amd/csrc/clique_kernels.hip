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
//     streamed as one flattened, coalesced sequence (a thread's element knows
//     its list by a forward cursor over the LDS segment offsets); each element
//     is looked up in L (binary search for one-word roots, a 16-bit position
//     hash otherwise) and, if present at position p, bit p of row i is set
//     (LDS / L2 atomic OR).  Row i is then the bitmap of L ∩ N+(L[i]), in
//     DAG order of the members.  Lists much longer than d (ratio switch) are
//     not streamed: every element of L is binary-searched in them instead
//     (galloping side of the merge/gallop choice).
//  2. Count.  k-cliques with source v0 are v0 plus a (k-1)-clique of the
//     DAG on L given by the rows.  TC needs no rows (it counts the matches of
//     step 1).  With d <= 64 every row is one 64-bit word and lane i counts
//     the cliques starting at i with AND + popcount.  With d > 64, the rows
//     are multi-word; for each i the wave compacts S = row i, and counts with
//     word-wise AND/popcount (k = 4), the compressed |S| x |S| successor
//     bitmap (k >= 5, |S| <= 64), fp4 matrix-core products (k = 5, wider S)
//     or register word windows (k >= 6).
// Counts are exact uint64 (the matrix-core products are exact integers
// below 2^24 in f32).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clique_dev.hpp"

namespace jk {

// ----------------------------------------------------------------------------
// Kernel.  G waves per root; WPB roots (groups) per block when G == 1.
// WMAX words per row => roots with d <= 64 * WMAX.  LROWS: rows in LDS,
// otherwise in a per-block slice of global scratch.
// Waves per SIMD the kernel's registers are held to.  These kernels are
// bound by latency rather than by any one unit, so on the tiers below the
// registers are capped at the occupancy their LDS allows (measured faster on
// the 5-clique tiers and the 4-clique 512-wide tier, also where they spill a
// few bytes a lane);
// elsewhere the compiler's own choice stands (already at the LDS or hardware
// limit, or the cap would spill).
template <int WMAX, int K, bool BYTES>
constexpr int min_waves() {
  if (BYTES) return 1;
  if (K == 4 && WMAX == 8) return 6;
  if (K == 5 && WMAX == 4) return 8;
  if (K == 5 && WMAX == 8) return 6;
  if (K == 6 && WMAX == 4) return 7;
  if (K == 7 && WMAX == 4) return 5;
  return 1;
}
template <int G, int WPB, int WMAX, int K, bool BYTES, bool LROWS>
__global__ void __launch_bounds__(64 * G * WPB) __attribute__((amdgpu_waves_per_eu(min_waves<WMAX, K, BYTES>())))
clique_kernel(CliqueArgs a) {
  using LY = Lay<G, WMAX, K, BYTES, LROWS>;
  constexpr int DM = LY::DM;
  constexpr int NT = 64 * G * WPB;
  __shared__ __attribute__((aligned(16))) unsigned char s_arena[WPB][LY::SLOT];
  __shared__ int32_t s_wtot[G + 1];
  __shared__ int64_t s_task;
  const int slot = (G == 1) ? (int)(threadIdx.x >> 6) : 0;
  const int ln = lane_id();
  unsigned char* ar = s_arena[slot];
  uint64_t* rowp = LROWS ? reinterpret_cast<uint64_t*>(ar + LY::R_OFF)
                         : (a.gscr + (int64_t)blockIdx.x * DM * WMAX);
  const int64_t* __restrict__ rp = a.rp;
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
      process_root<G, WMAX, K, BYTES, LROWS>(rp, a.ci, a.ord, b0, d, a.ratio, ar, rowp, s_wtot, acc);
    }
  }
  block_add<NT>(acc, a.out);
}

// ---------------------------------------------------------------------------
// Split form of the 1024-wide tier: the build runs with 16 waves per root
// (twice the waves that the count's registers and per-wave lists allow) and
// leaves each root's rows in a global slot; the count kernel reloads them.
// Task t (of a.ids) uses slot t of a.gscr, ROWW words each.
constexpr int SPLIT_WMAX = 16;
constexpr size_t ROWW = (size_t)64 * SPLIT_WMAX * SPLIT_WMAX;

template <int K>
__global__ void __launch_bounds__(64 * 16) clique_build_kernel(CliqueArgs a) {
  constexpr int G = 16;
  using LY = Lay<G, SPLIT_WMAX, K, false, true, false>;
  __shared__ __attribute__((aligned(16))) unsigned char s_arena[LY::SLOT];
  __shared__ int32_t s_wtot[G + 1];
  __shared__ int64_t s_task;
  uint64_t* rowp = reinterpret_cast<uint64_t*>(s_arena + LY::R_OFF);
  uint64_t acc = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = (int64_t)atomicAdd(a.qctr, 1u);
    __syncthreads();
    const int64_t task = s_task;
    __syncthreads();
    if (task >= a.ntasks) break;
    const int32_t v0 = a.ids[task];
    const int64_t b0 = a.rp[v0];
    const int d = (int)(a.rp[v0 + 1] - b0);
    if (d < a.dmin || d > a.dmax || d > LY::DM) continue;
    process_root<G, SPLIT_WMAX, K, false, true, 1>(a.rp, a.ci, a.ord, b0, d, a.ratio, s_arena, rowp, s_wtot, acc);
    const int n = d * RowLayout((d + 63) >> 6).S;
    uint64_t* dst = a.gscr + (size_t)task * ROWW;
    for (int t = threadIdx.x; t < n; t += 64 * G) dst[t] = rowp[t];
    __syncthreads();
  }
}

// Waves per root of the count kernel: 12 (three per SIMD, all the LDS
// allows beside the rows) where the count fits the registers of three waves;
// the 7-clique count would spill there, so it keeps 8.
template <int K>
constexpr int split_count_waves() { return (K == 7) ? 8 : 12; }

template <int K>
__global__ void __launch_bounds__(64 * split_count_waves<K>()) clique_count_kernel(CliqueArgs a) {
  constexpr int G = split_count_waves<K>();
  using LY = Lay<G, SPLIT_WMAX, K, false, true, true, false>;
  __shared__ __attribute__((aligned(16))) unsigned char s_arena[LY::SLOT];
  __shared__ int32_t s_wtot[G + 1];
  __shared__ int64_t s_task;
  uint64_t* rowp = reinterpret_cast<uint64_t*>(s_arena + LY::R_OFF);
  uint64_t acc = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = (int64_t)atomicAdd(a.qctr, 1u);
    __syncthreads();
    const int64_t task = s_task;
    __syncthreads();
    if (task >= a.ntasks) break;
    const int32_t v0 = a.ids[task];
    const int64_t b0 = a.rp[v0];
    const int d = (int)(a.rp[v0 + 1] - b0);
    if (d < a.dmin || d > a.dmax || d > LY::DM) continue;
    const int n = d * RowLayout((d + 63) >> 6).S;
    const uint64_t* src = a.gscr + (size_t)task * ROWW;
    for (int t = threadIdx.x; t < n; t += 64 * G) rowp[t] = __builtin_nontemporal_load(src + t);
    if (threadIdx.x == 0) s_wtot[G] = 0;
    __syncthreads();
    process_root<G, SPLIT_WMAX, K, false, true, 2>(a.rp, a.ci, a.ord, b0, d, a.ratio, s_arena, rowp, s_wtot, acc);
  }
  block_add<64 * G>(acc, a.out);
}

template <int K>
static hipError_t launch_split_k(bool build, const CliqueArgs& a, int grid, hipStream_t st) {
  if (build) hipLaunchKernelGGL(clique_build_kernel<K>, dim3(grid), dim3(64 * 16), 0, st, a);
  else hipLaunchKernelGGL(clique_count_kernel<K>, dim3(grid), dim3(64 * split_count_waves<K>()), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_clique_split(int k, bool build, const CliqueArgs& a, int grid, hipStream_t st) {
  switch (k) {
    case 4: return launch_split_k<4>(build, a, grid, st);
    case 5: return launch_split_k<5>(build, a, grid, st);
    case 6: return launch_split_k<6>(build, a, grid, st);
    case 7: return launch_split_k<7>(build, a, grid, st);
  }
  return hipErrorInvalidValue;
}

size_t split_slot_words() { return ROWW; }

// ---------------------------------------------------------------------------
// Split form of the 2048-wide tier for 5- to 7-cliques.  The fused kernel
// counts a root with one block of 8 waves (two per SIMD, the registers of
// the 2x2 matrix-core tiles), so the roots of a query run about one per CU
// and the densest ones leave the device waiting.  Here a build kernel with
// 16 waves per root leaves each root's rows (512 KiB) in a global slot, and
// the count kernel lets a.rep blocks take the rows of one root from a shared
// global row counter.  (4-cliques keep the fused kernel: their whole-root
// product is not split by rows.)
constexpr int GLB_WMAX = 32;
constexpr size_t GSLOT = (size_t)64 * GLB_WMAX * GLB_WMAX;

template <int K>
__global__ void __launch_bounds__(64 * 16) clique_glb_build_kernel(CliqueArgs a) {
  constexpr int G = 16;
  using LY = Lay<G, GLB_WMAX, K, false, false, false>;
  __shared__ __attribute__((aligned(16))) unsigned char s_arena[LY::SLOT];
  __shared__ int32_t s_wtot[G + 1];
  __shared__ int64_t s_task;
  uint64_t acc = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = (int64_t)atomicAdd(a.qctr, 1u);
    __syncthreads();
    const int64_t task = s_task;
    __syncthreads();
    if (task >= a.ntasks) break;
    const int32_t v0 = a.ids[task];
    const int64_t b0 = a.rp[v0];
    const int d = (int)(a.rp[v0 + 1] - b0);
    if (d < a.dmin || d > a.dmax || d > LY::DM) continue;
    process_root<G, GLB_WMAX, K, false, false, 1>(a.rp, a.ci, a.ord, b0, d, a.ratio, s_arena,
                                                  a.gscr + (size_t)task * GSLOT, s_wtot, acc);
  }
}

template <int K>
__global__ void __launch_bounds__(64 * 8) clique_glb_count_kernel(CliqueArgs a) {
  constexpr int G = 8;
  using LY = Lay<G, GLB_WMAX, K, false, false, true, false>;
  __shared__ __attribute__((aligned(16))) unsigned char s_arena[LY::SLOT];
  __shared__ int32_t s_wtot[G + 1];
  __shared__ int64_t s_item;
  uint64_t acc = 0;
  // Workgroups are dealt to the 8 XCDs round robin, so block b runs on XCD
  // b % 8.  Queue q (counter a.rowctr[q - 8]) holds the roots q, q + 8, ...,
  // each as rep items that start together: the blocks sharing a root's
  // 512 KiB of rows are on one XCD and read them through its L2.  A block
  // drains its own XCD's queue, then helps the others.
  uint32_t* const qx = a.rowctr - 8;
  const int xcd = (int)(blockIdx.x & 7);
  for (int s = 0; s < 8; ++s) {
    const int q = (xcd + s) & 7;
    const int64_t nq = a.ntasks > q ? ((a.ntasks - 1 - q) / 8 + 1) * a.rep : 0;
    for (;;) {
      if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(qx + q, 1u);
      __syncthreads();
      const int64_t item = s_item;
      __syncthreads();
      if (item >= nq) break;
      const int64_t task = (item / a.rep) * 8 + q;
      const int32_t v0 = a.ids[task];
      const int64_t b0 = a.rp[v0];
      const int d = (int)(a.rp[v0 + 1] - b0);
      if (d < a.dmin || d > a.dmax || d > LY::DM) continue;
      process_root<G, GLB_WMAX, K, false, false, 2>(a.rp, a.ci, a.ord, b0, d, a.ratio, s_arena,
                                                    a.gscr + (size_t)task * GSLOT, s_wtot, acc, a.rowctr + task);
    }
  }
  block_add<64 * G>(acc, a.out);
}

template <int K>
static hipError_t launch_glb_split_k(bool build, const CliqueArgs& a, int grid, hipStream_t st) {
  if (build) hipLaunchKernelGGL(clique_glb_build_kernel<K>, dim3(grid), dim3(64 * 16), 0, st, a);
  else hipLaunchKernelGGL(clique_glb_count_kernel<K>, dim3(grid), dim3(64 * 8), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_clique_glb_split(int k, bool build, const CliqueArgs& a, int grid, hipStream_t st) {
  switch (k) {
    case 5: return launch_glb_split_k<5>(build, a, grid, st);
    case 6: return launch_glb_split_k<6>(build, a, grid, st);
    case 7: return launch_glb_split_k<7>(build, a, grid, st);
  }
  return hipErrorInvalidValue;
}

size_t glb_slot_words() { return GSLOT; }

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
    case TIER_LDS8:   return launch_one<8, 1, 8, K, BYTES, true>(a, grid, st);
    case TIER_LDS16:  return launch_one<8, 1, 16, K, BYTES, true>(a, grid, st);
    case TIER_GLB32:  return launch_one<8, 1, 32, K, BYTES, false>(a, grid, st);
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
    case TIER_LDS16: return 1024;
    case TIER_GLB32: return 2048;
  }
  return 0;
}

size_t tier_scratch_words(int tier) {
  return tier == TIER_GLB32 ? (size_t)2048 * 32 : 0;
}

}  // namespace jk
