// batch.hip — many (graph, pattern, root range) queries per call.
#include <hip/hip_runtime.h>

#include "jgpm_device.hpp"

extern "C" int jgpm_count_batch(jgpm_ctx_t ctx, const jgpm_query* q, int n, uint64_t* counts, double* kernel_ms) {
  if (!ctx || (n > 0 && (!q || !counts)) || n < 0) return JGPM_E_INVALID;
  double total = 0;
  for (int i = 0; i < n; ++i) {
    double ms = 0;
    int st = jgpm_count(q[i].graph, q[i].pattern, q[i].root_begin, q[i].root_end, &counts[i], &ms);
    if (st) return st;
    total += ms;
  }
  if (kernel_ms) *kernel_ms = total;
  return JGPM_OK;
}
