// clique_huge.hip — k-cliques (TC, P13..P16) of the roots wider than the
// widest bitmap-row tier (DAG out-degree above 2048).
//
// Such roots are rare (none in R-MAT up to scale 24, whose DAG out-degree is
// at most 1775: the degree orientation bounds it by sqrt(2E)), so they run
// the level nest of the reference's k-clique plan directly
// (src/config.h:39-58, src/gpu_kernels_codegen.cuh:261-410; message-passing
// form src/message_passing_gpu_kernels.cuh:1208-1537): T_1 = N+(v0),
// T_{i+1} = T_i ∩ N+(v_i) materialised in global scratch, the last level
// counted as |T ∩ N+(v)|.  The k-cliques found are exactly those whose DAG
// source is v0, the attribution of every other tier.  One block of HW waves
// per root (roots dealt to blocks by a stride); the waves take the members v1
// of T_1 round-robin; the set operations are the wave-wide ones of the
// symmetric kernels (ascending lists, LDS-staged short lists).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clique_dev.hpp"
#include "kernels.hpp"
#include "sym_dev.hpp"

namespace jk {

using namespace symd;

namespace {

constexpr int HW = 4;  // waves per block

// Number of R-cliques of the DAG inside the ascending list T (n members);
// lists below T go to scr (R - 2 lists of stride ints).
template <int R>
__device__ uint64_t rcliques(const int64_t* rp, const int32_t* ci, Wave& w, const int32_t* T, int n, int32_t* scr,
                             int64_t stride) {
  if constexpr (R == 1) {
    return (uint64_t)n;
  } else if constexpr (R == 2) {
    uint64_t s = 0;
    for (int i = 0; i < n; ++i) {
      const int32_t v = T[i];
      s += (uint64_t)w_icount(w, T, n, ci + rp[v], (int)(rp[v + 1] - rp[v]), INT32_MAX);
    }
    return s;
  } else {
    uint64_t s = 0;
    for (int i = 0; i < n; ++i) {
      const int32_t v = T[i];
      const int m = w_imat(w, T, n, ci + rp[v], (int)(rp[v + 1] - rp[v]), INT32_MAX, scr, nullptr, nullptr);
      if (m >= R - 1) s += rcliques<R - 1>(rp, ci, w, scr, m, scr + stride, stride);
    }
    return s;
  }
}

template <int K>
__global__ void __launch_bounds__(64 * HW) clique_huge_kernel(HugeArgs a) {
  __shared__ int32_t s_stage[HW][STAGE];
  __shared__ uint64_t s_part[HW];
  const int wv = threadIdx.x >> 6;
  Wave w{s_stage[wv]};
  int32_t* scr = a.scr + ((int64_t)blockIdx.x * HW + wv) * (K >= 4 ? K - 3 : 1) * a.stride;
  uint64_t acc = 0;
  for (int64_t t = blockIdx.x; t < a.ntasks; t += gridDim.x) {
    const int32_t v0 = a.ids[t];
    const int32_t* T1 = a.ci + a.rp[v0];
    const int d = (int)(a.rp[v0 + 1] - a.rp[v0]);
    for (int i = wv; i < d; i += HW) {
      const int32_t v1 = T1[i];
      const int32_t* N1 = a.ci + a.rp[v1];
      const int d1 = (int)(a.rp[v1 + 1] - a.rp[v1]);
      if constexpr (K == 3) {
        acc += (uint64_t)w_icount(w, T1, d, N1, d1, INT32_MAX);
      } else {
        const int m = w_imat(w, T1, d, N1, d1, INT32_MAX, scr, nullptr, nullptr);
        if (m >= K - 2) acc += rcliques<K - 2>(a.rp, a.ci, w, scr, m, scr + a.stride, a.stride);
      }
    }
  }
  if (lane() == 0) s_part[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int k = 0; k < HW; ++k) s += s_part[k];
    if (s) atomicAdd((unsigned long long*)a.out, (unsigned long long)s);
  }
}

}  // namespace

int huge_lists_per_block(int k) { return HW * (k >= 4 ? k - 3 : 1); }

hipError_t launch_clique_huge(int k, const HugeArgs& a, int grid, hipStream_t st) {
  switch (k) {
    case 3: hipLaunchKernelGGL(clique_huge_kernel<3>, dim3(grid), dim3(64 * HW), 0, st, a); break;
    case 4: hipLaunchKernelGGL(clique_huge_kernel<4>, dim3(grid), dim3(64 * HW), 0, st, a); break;
    case 5: hipLaunchKernelGGL(clique_huge_kernel<5>, dim3(grid), dim3(64 * HW), 0, st, a); break;
    case 6: hipLaunchKernelGGL(clique_huge_kernel<6>, dim3(grid), dim3(64 * HW), 0, st, a); break;
    case 7: hipLaunchKernelGGL(clique_huge_kernel<7>, dim3(grid), dim3(64 * HW), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace jk
