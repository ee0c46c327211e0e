// Device-side state of the C ABI objects (context and resident graph).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "jgpm.h"
#include "jgpm_internal.hpp"
#include "kernels.hpp"

// Accumulator / queue-counter slots per context (one query uses a few).
constexpr int CTX_SLOTS = 64;

struct DevCSR {
  int64_t* rp = nullptr;
  int32_t* ci = nullptr;
  int32_t nv = 0;
  int64_t nnz = 0;
  int32_t max_deg = 0;
};

struct jgpm_ctx_s {
  int dev = 0;
  hipStream_t st = nullptr;
  int num_cu = 0;
  int lds = 0;
  int64_t hbm = 0;
  char name[128] = {0};
  jgpm_tuning tune{};
  uint64_t* d_out = nullptr;  // CTX_SLOTS accumulators
  uint32_t* d_q = nullptr;    // CTX_SLOTS work-queue counters
  uint64_t* d_scr = nullptr;  // row scratch of the global-row tier
  size_t scr_words = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::mutex mu;
};

struct jgpm_graph_s {
  jgpm_ctx_s* ctx = nullptr;
  bool has_sym = false, has_dag = false;
  DevCSR sym, dag;
  std::vector<int32_t> tier_h[jk::NUM_TIERS];  // root ids per block tier (ascending)
  int32_t* tier_d[jk::NUM_TIERS] = {nullptr};
  std::vector<int32_t> huge_h;                 // roots beyond the largest tier
  int64_t bytes = 0;
};

namespace jgpm {
int ctx_reset_counters(jgpm_ctx_s* c, hipStream_t st);
int ensure_scratch(jgpm_ctx_s* c, size_t words);
int clique_k(int pattern);
int enqueue_clique(jgpm_graph_s* g, int k, bool bytes, int64_t rb, int64_t re, int out_slot, hipStream_t st);
// Enqueue every kernel of `pattern` for roots [rb, re); results land in
// accumulator slots base_slot.. (see finish_pattern).
int enqueue_pattern(jgpm_graph_s* g, int pattern, int64_t rb, int64_t re, int base_slot, hipStream_t st);
// Number of accumulator slots a pattern uses.
int pattern_slots(int pattern);
// Combine the accumulator slots of a finished query into its count.
uint64_t finish_pattern(int pattern, const uint64_t* slots);
}  // namespace jgpm
