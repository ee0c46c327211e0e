// vs_driver.cc — C entry points over a VertexSet implementation, compiled
// twice by oracle/Makefile: against the reference's own VertexSet
// (include/VertexSet.h, common/VertexSet.cc under the reference tree ->
// oracle/_ref/libref_vs.so) and against this repository's drop-in
// (include/VertexSet.h, jupiter_amd/csrc/VertexSet.cc -> oracle/_build/libmy_vs.so).
// TEST INFRASTRUCTURE ONLY: tests/test_vertexset.py compares the two.
#include <cstdint>
#include <cstring>

#include "VertexSet.h"

extern "C" {

void vs_set_max_degree(int32_t m) {
  if (m > VertexSet::MAX_DEGREE) VertexSet::MAX_DEGREE = m;
}

// op: 0 a&b, 1 intersect(upper), 2 a-b, 3 difference(upper), 4 bounded(upper)
int32_t vs_set_op(int op, int32_t* a, int32_t na, int32_t avid, int32_t* b, int32_t nb, int32_t bvid, int32_t upper,
                  int32_t* out) {
  VertexSet A(a, na, avid), B(b, nb, bvid);
  if (op == 0) {
    VertexSet r = A & B;
    std::memcpy(out, r.data(), sizeof(int32_t) * r.size());
    return r.size();
  }
  if (op == 1) {
    VertexSet r = A.intersect(B, upper);
    std::memcpy(out, r.data(), sizeof(int32_t) * r.size());
    return r.size();
  }
  if (op == 2) {
    VertexSet r = A - B;
    std::memcpy(out, r.data(), sizeof(int32_t) * r.size());
    return r.size();
  }
  if (op == 3) {
    VertexSet r = A.difference(B, upper);
    std::memcpy(out, r.data(), sizeof(int32_t) * r.size());
    return r.size();
  }
  VertexSet r = A.bounded(upper);
  std::memcpy(out, r.data(), sizeof(int32_t) * r.size());
  return r.size();
}

// op: 0 intersection_num, 1 intersection_num(upper), 2 except(anc), 3 except(anc, anc2),
//     4 bound_except(upper, anc), 5 difference_num, 6 difference_num(upper)
uint64_t vs_count_op(int op, int32_t* a, int32_t na, int32_t avid, int32_t* b, int32_t nb, int32_t bvid,
                     int32_t upper, int32_t anc, int32_t anc2) {
  VertexSet A(a, na, avid), B(b, nb, bvid);
  switch (op) {
    case 0: return intersection_num(A, B);
    case 1: return intersection_num(A, B, upper);
    case 2: return intersection_num_except(A, B, anc);
    case 3: return intersection_num_except(A, B, anc, anc2);
    case 4: return intersection_num_bound_except(A, B, upper, anc);
    case 5: return difference_num(A, B);
    case 6: return difference_num(A, B, upper);
  }
  return 0;
}
}
