// plan_kernels.hip — executor of compiled pattern plans (plan.cpp): counts
// any connected pattern of up to 7 vertices on the symmetric CSR.
//
// One wave owns one root v_0 (a vertex of the root range matched at level 0)
// and walks the plan depth first.  Level l's candidates are the intersection
// of the neighbour lists of its constraining earlier levels, starting from
// the shortest list, trimmed by binary search to the id range that the
// symmetry-breaking restrictions leave, then filtered: for vertex-induced
// plans the neighbours of the non-adjacent earlier levels are removed by a
// wave-wide set difference (w_idiff), and earlier vertices are removed
// (injectivity).  The wave-wide set operations are the ones
// of the symmetric kernels (sym_dev.hpp).  The last level is counted without
// materialising when it has at most two constraining lists: |A ∩ B| minus the
// earlier vertices it contains.  Levels are a compile-time recursion, so the
// per-level state stays in registers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "jgpm_device.hpp"
#include "kernels.hpp"
#include "sym_dev.hpp"

namespace jk {

using namespace symd;

namespace {

__device__ __forceinline__ int32_t uni32(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

struct PCtx {
  const int64_t* rp;
  const int32_t* ci;
  int32_t* scr;    // (k + 1) lists of `stride` ints: one per level, one temporary
  int64_t stride;
  int k;
  __device__ const int32_t* N(int32_t v) const { return ci + rp[v]; }
  __device__ int d(int32_t v) const { return (int)(rp[v + 1] - rp[v]); }
};

// In place: drop the neighbours of every non-adjacent earlier level (nadj),
// one set difference X − N(v_p) per such level, then every element equal to
// an earlier level's vertex (injectivity).
template <int L>
__device__ int w_filter(const PCtx& c, Wave& w, int32_t* X, int n, const int32_t (&v)[PLAN_MAXK], uint32_t nadj) {
#pragma unroll
  for (int p = 0; p < L; ++p)
    if ((nadj >> p) & 1u) n = w_idiff(w, X, n, c.N(v[p]), c.d(v[p]), INT32_MAX, X);
  int m = 0;
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane();
    bool keep = false;
    int32_t x = 0;
    if (i < n) {
      x = X[i];
      keep = true;
#pragma unroll
      for (int p = 0; p < L; ++p) keep = keep && x != v[p];
    }
    const uint64_t b = __ballot(keep);
    if (keep)
      X[m + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = x;
    m += __popcll(b);
  }
  __syncwarp();
  return m;
}

template <int L>
__device__ uint64_t plan_level(const PCtx& c, Wave& w, const Plan& P, int32_t (&v)[PLAN_MAXK]) {
  // ids the restrictions leave for v_L: (lo, hi), both exclusive
  int32_t lo = -1, hi = INT32_MAX;
#pragma unroll
  for (int p = 0; p < L; ++p) {
    if ((P.gt[L] >> p) & 1u) lo = max(lo, v[p]);
    if ((P.lt[L] >> p) & 1u) hi = min(hi, v[p]);
  }
  // the shortest constraining neighbour list, trimmed to (lo, hi)
  int p0 = 0, na = INT32_MAX;
  const int32_t* A = nullptr;
#pragma unroll
  for (int p = 0; p < L; ++p) {
    if ((P.conn[L] >> p) & 1u) {
      const int dp = c.d(v[p]);
      if (dp < na) {
        na = dp;
        p0 = p;
        A = c.N(v[p]);
      }
    }
  }
  {
    const int a0 = lo >= 0 ? lb(A, na, lo + 1) : 0;
    const int a1 = hi != INT32_MAX ? lb(A, na, hi) : na;
    A += a0;
    na = max(0, a1 - a0);
  }
  const uint32_t others = P.conn[L] & ~(1u << p0);
  const bool last = L == P.k - 1;
  const uint32_t nadj = P.induced ? P.nconn[L] : 0u;
  if (!nadj && __builtin_popcount(others) <= 1) {
    if (last) {
      const int32_t* B = nullptr;
      int nb = 0;
#pragma unroll
      for (int p = 0; p < L; ++p)
        if ((others >> p) & 1u) {
          B = c.N(v[p]);
          nb = c.d(v[p]);
        }
      int64_t cnt = B ? w_icount(w, A, na, B, nb, INT32_MAX) : (int64_t)na;
#pragma unroll
      for (int p = 0; p < L; ++p)  // earlier vertices the count includes
        if (in_sorted(A, na, v[p]) && (!B || in_sorted(B, nb, v[p]))) --cnt;
      return (uint64_t)cnt;
    }
    if (!others) {  // iterate the list itself, skipping earlier vertices
      uint64_t t = 0;
      if constexpr (L + 1 < PLAN_MAXK) {
        for (int i = 0; i < na; ++i) {
          const int32_t x = uni32(A[i]);
          bool dup = false;
#pragma unroll
          for (int p = 0; p < L; ++p) dup = dup || x == v[p];
          if (dup) continue;
          v[L] = x;
          t += plan_level<L + 1>(c, w, P, v);
        }
      }
      return t;
    }
  }
  // materialise the candidates of level L in its own list
  int32_t* out = c.scr + (int64_t)L * c.stride;
  int32_t* tmp = c.scr + (int64_t)c.k * c.stride;
  const int m = __builtin_popcount(others);
  int nc = na;
  if (m == 0) {
    for (int i = lane(); i < na; i += 64) out[i] = A[i];
    __syncwarp();
  } else {
    int32_t* dst = (m & 1) ? out : tmp;  // so that the last intersection lands in out
    const int32_t* cur = A;
#pragma unroll
    for (int p = 0; p < L; ++p) {
      if ((others >> p) & 1u) {
        nc = w_imat(w, cur, nc, c.N(v[p]), c.d(v[p]), INT32_MAX, dst, nullptr, nullptr);
        cur = dst;
        dst = dst == out ? tmp : out;
      }
    }
  }
  nc = w_filter<L>(c, w, out, nc, v, nadj);
  if (last) return (uint64_t)nc;
  uint64_t t = 0;
  if constexpr (L + 1 < PLAN_MAXK) {
    for (int i = 0; i < nc; ++i) {
      v[L] = uni32(out[i]);
      t += plan_level<L + 1>(c, w, P, v);
    }
  }
  return t;
}

__global__ void __launch_bounds__(64 * PLAN_WPB) plan_kernel(PlanArgs a) {
  __shared__ int32_t s_stage[PLAN_WPB][STAGE];
  const int wv = threadIdx.x >> 6;
  Wave w{s_stage[wv]};
  const Plan& P = a.plan;
  const PCtx c{a.rp, a.ci, a.scr + ((int64_t)blockIdx.x * PLAN_WPB + wv) * (P.k + 1) * a.stride, a.stride, P.k};
  uint64_t acc = 0;
  for (;;) {
    int64_t t = 0;
    if (lane() == 0) t = (int64_t)atomicAdd(a.qctr, 1ull);
    const int64_t r = a.rb + __shfl(t, 0);
    if (r >= a.re) break;
    int32_t v[PLAN_MAXK] = {0};
    v[0] = (int32_t)r;
    acc += plan_level<1>(c, w, P, v);
  }
  if (lane() == 0 && acc) atomicAdd((unsigned long long*)a.out, (unsigned long long)acc);
}

}  // namespace

hipError_t launch_plan(const PlanArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(plan_kernel, dim3(grid), dim3(64 * PLAN_WPB), 0, st, a);
  return hipGetLastError();
}

}  // namespace jk

namespace jgpm {
const jk::Plan* plan_of(const jgpm_plan_s* p);
}

extern "C" int jgpm_count_plan(jgpm_graph_t g, jgpm_plan_t p, int64_t rb, int64_t re, uint64_t* count,
                               double* kernel_ms) {
  if (!g || !p || !count) return JGPM_E_INVALID;
  if (!g->has_sym) return JGPM_E_GRAPHKIND;
  const jk::Plan* P = jgpm::plan_of(p);
  const int64_t nv = g->sym.nv;
  if (re < 0) re = nv;
  if (rb < 0 || rb > re || re > nv) return JGPM_E_INVALID;
  jgpm_ctx_s* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->dev) != hipSuccess) return JGPM_E_HIP;
  // per-wave lists: (k + 1) of max-degree length, within a 2 GiB scratch budget
  const int64_t stride = std::max<int64_t>(1, g->sym.max_deg);
  const int64_t per_wave = (int64_t)(P->k + 1) * stride;
  const int64_t budget = (int64_t)1 << 29;  // ints
  const int64_t waves = std::max<int64_t>(jk::PLAN_WPB, std::min<int64_t>(budget / per_wave, (int64_t)c->num_cu * 16));
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(waves / jk::PLAN_WPB, std::max<int64_t>(1, re - rb)));
  const size_t ints = (size_t)grid * jk::PLAN_WPB * (size_t)per_wave;
  int st = jgpm::ensure_scratch(c, ints / 2 + 1);
  if (st) return st;
  st = jgpm::ctx_reset_counters(c, c->st);
  if (st) return st;
  jk::PlanArgs a{};
  a.rp = g->sym.rp;
  a.ci = g->sym.ci;
  a.rb = rb;
  a.re = re;
  a.qctr = c->d_q64;
  a.scr = reinterpret_cast<int32_t*>(c->d_scr);
  a.stride = stride;
  a.out = c->d_out;
  a.plan = *P;
  c->nprof = 0;
  uint64_t h = 0;
  if (hipEventRecord(c->e0, c->st) != hipSuccess) return JGPM_E_HIP;
  if (re > rb && jk::launch_plan(a, grid, c->st) != hipSuccess) return JGPM_E_HIP;
  if (hipEventRecord(c->e1, c->st) != hipSuccess ||
      hipMemcpyAsync(&h, c->d_out, sizeof(uint64_t), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return JGPM_E_HIP;
  *count = h;
  if (kernel_ms) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->e0, c->e1);
    *kernel_ms = ms;
  }
  return JGPM_OK;
}
