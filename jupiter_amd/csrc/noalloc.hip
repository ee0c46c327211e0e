// noalloc.hip — jgpm_count_noalloc: the clique counts (TC, P13..P16) of
// jgpm_count without any device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  No counterpart in the reference.
//
// The query's device scratch lives in one file-scope array per GPU:
//   words [0, 64)      accumulators (the count lands in word 0)
//   words [64, 128)    work-queue counters (32-bit, one per clique tier)
//   words [256, ...)   bitmap rows of the 2048-wide tier, one slice of
//                      jk::tier_scratch_words(TIER_GLB32) words per block
// The array is device memory the code object holds from load until the
// process ends.  Its address is taken once per device with
// hipGetSymbolAddress (noalloc_setup, called by jgpm_init); host code passes
// only that address to kernels, never the array's name.  Every call on a
// GPU shares the array and leaves its contents behind: the accumulators and
// counters are cleared by fill kernels at the start of each call, the rows
// are written before they are read.  Calls on one GPU that overlap in time
// (two contexts on the same device, two threads with two contexts) race on
// the array; calls through one context are serialised by its mutex.
#include <hip/hip_runtime.h>

#include <mutex>

#include "jgpm_device.hpp"
#include "jgpm_internal.hpp"
#include "kernels.hpp"

namespace {

constexpr size_t kScratchWords = (size_t)4 << 20;  // 32 MiB per GPU
constexpr size_t kOutWord = 0, kQueueWord = 64, kRowsWord = 256;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDev = 64;
std::mutex g_mu;
uint64_t* g_ptr[kMaxDev] = {nullptr};

__global__ void k_noalloc_clear(uint64_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

namespace jgpm {

// Device address of this GPU's scratch array, taken once per device with the
// device current.  Called by jgpm_init; later calls return the cached value.
int noalloc_setup(int dev, uint64_t** out) {
  if (dev < 0 || dev >= kMaxDev) return JGPM_E_INVALID;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != dev) return JGPM_E_STATE;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return JGPM_E_HIP;
    g_ptr[dev] = static_cast<uint64_t*>(p);
  }
  if (out) *out = g_ptr[dev];
  return JGPM_OK;
}

}  // namespace jgpm

extern "C" int jgpm_count_noalloc(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end,
                                  uint64_t* count) {
  if (!g || !count) return JGPM_E_INVALID;
  const int k = jgpm::clique_k(pattern);
  if (!k) return JGPM_E_PATTERN;  // only the clique engine runs without allocation
  if (!g->has_dag) return JGPM_E_GRAPHKIND;
  const int32_t nv = g->dag.nv;
  if (root_end < 0) root_end = nv;
  if (root_begin < 0 || root_begin > root_end || root_end > nv) return JGPM_E_INVALID;
  jgpm_ctx_s* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->dev) != hipSuccess) return JGPM_E_HIP;
  uint64_t* base = nullptr;
  int st = jgpm::noalloc_setup(c->dev, &base);
  if (st) return st;
  // one 2048-wide root needs one slice; the tier runs with as many blocks as
  // the array holds slices (at least one), so no root range exceeds it
  const size_t slice = jk::tier_scratch_words(jk::TIER_GLB32);
  const size_t rows_words = kScratchWords - kRowsWord;
  if (slice > rows_words) return JGPM_E_BUDGET;
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(256), 0, c->st, base, (int)kRowsWord);
  if (hipGetLastError() != hipSuccess) return JGPM_E_HIP;
  const bool prof = c->prof_on;
  c->prof_on = false;  // the profile events belong to jgpm_count's queries
  st = jgpm::enqueue_clique(g, k, false, root_begin, root_end, 0, c->st, false, 0, base + kOutWord,
                            reinterpret_cast<unsigned*>(base + kQueueWord), base + kRowsWord, rows_words);
  c->prof_on = prof;
  if (st) {
    (void)hipStreamSynchronize(c->st);
    return st;
  }
  uint64_t h = 0;
  if (hipMemcpyAsync(&h, base + kOutWord, sizeof(h), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return JGPM_E_HIP;
  *count = h;
  return JGPM_OK;
}
