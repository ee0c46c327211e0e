// plan.hpp — compiled pattern plan shared by the host compiler (plan.cpp) and
// the device executor (plan_kernels.hip).  Plain data, no HIP types.
#pragma once
#include <stdint.h>

namespace jk {

// Level l of the matching order is matched to a vertex v_l; bit p of a mask
// refers to the earlier level p < l.
constexpr int PLAN_MAXK = 7;
struct Plan {
  int32_t k;
  int32_t induced;               // 1: vertex-induced (non-adjacent pattern pairs must be non-adjacent)
  uint32_t conn[PLAN_MAXK];      // v_l must be adjacent to v_p for p in conn[l] (never empty for l > 0)
  uint32_t nconn[PLAN_MAXK];     // induced only: v_l must not be adjacent to v_p for p in nconn[l]
  uint32_t gt[PLAN_MAXK];        // symmetry breaking: v_l > v_p for p in gt[l]
  uint32_t lt[PLAN_MAXK];        // symmetry breaking: v_l < v_p for p in lt[l]
};

}  // namespace jk
