// Device kernel entry points of libjgpm (host-callable launchers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jk {

// Root tiers of the clique engine, by DAG out-degree d of the root.
enum {
  TIER_WAVE = 0,   // d <= 64    one wave per root, one-word rows in LDS
  TIER_LDS2 = 1,   // d <= 128   block per root, rows in LDS
  TIER_LDS4 = 2,   // d <= 256
  TIER_LDS8 = 3,   // d <= 512
  TIER_GLB32 = 4,  // d <= 2048  block per root, rows in a global scratch slice
  NUM_TIERS = 5
};

struct CliqueArgs {
  const int64_t* rp;     // DAG row pointers
  const int32_t* ci;     // DAG column indices
  const int32_t* ids;    // task -> root vertex (NULL: root = rb + task)
  int64_t rb;
  int64_t ntasks;
  int32_t dmin, dmax;    // roots outside [dmin, dmax] are skipped by this launch
  int32_t ratio;         // a list longer than ratio * d is binary-searched, not streamed
  int32_t chunk;         // roots taken per queue grab (wave tier)
  uint32_t* qctr;        // work-queue counter (zero before launch)
  uint64_t* gscr;        // global row scratch (TIER_GLB32)
  uint64_t* out;         // uint64 accumulator
};

hipError_t launch_clique(int k, bool bytes, int tier, const CliqueArgs& a, int grid, hipStream_t st);
int tier_max_deg(int tier);
size_t tier_scratch_words(int tier);

}  // namespace jk
