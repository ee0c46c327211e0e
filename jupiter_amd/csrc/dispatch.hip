// dispatch.hip — pattern id -> kernels, and the combination of partial
// accumulators into the final count.
//
// The reference chooses kernels by string compare and asserts on an unknown
// name (src/jupiter_gpm.cu:391-433,435-873); here an id selects an enqueue
// routine and an unknown id is JGPM_E_PATTERN.
#include <hip/hip_runtime.h>

#include "jgpm_device.hpp"

namespace jgpm {

int pattern_slots(int pattern) {
  switch (pattern) {
    case JGPM_I_C4: case JGPM_I_DIAMOND: return 3;
    case JGPM_I_TAILTRI: return 3;
    case JGPM_I_STAR3: return 4;
  }
  return 1;
}

int enqueue_pattern(jgpm_graph_s* g, int pattern, int64_t rb, int64_t re, int s, hipStream_t st) {
  int k = clique_k(pattern);
  if (k) {
    if (!g->has_dag) return JGPM_E_GRAPHKIND;
    return enqueue_clique(g, k, false, rb, re, s, st);
  }
  return JGPM_E_PATTERN;
}

uint64_t finish_pattern(int pattern, const uint64_t* h) {
  (void)pattern;
  return h[0];
}

}  // namespace jgpm
