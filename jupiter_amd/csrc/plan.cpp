// plan.cpp — compiler from a small pattern graph to a matching plan for the
// general executor (plan_kernels.hip).
//
// The reference hand-writes one kernel per pattern id and leaves a general
// generator as a TODO (src/gpu_kernels_codegen.cuh:3-4; plan levels in
// src/config.h:39-58).  Here any connected pattern of up to 7 vertices is
// compiled on the host into:
//   * a matching order: the pattern vertex of largest degree first, then
//     repeatedly the unmatched vertex with most matched neighbours (ties:
//     larger degree, then smaller id), so every later level has an earlier
//     neighbour and its candidates are an intersection of neighbour lists;
//   * for every level, the earlier levels it must be adjacent to and, for
//     vertex-induced counting, the ones it must not be adjacent to;
//   * symmetry-breaking restrictions v_a < v_b from the automorphism group
//     (stabiliser chain: take the first level, in matching order, whose
//     orbit under the remaining group is non-trivial, require it to map to
//     the smallest id of its orbit, continue with its stabiliser).  Then each
//     occurrence of the pattern is matched exactly once, so the number of
//     matches is the number of (non-induced or induced) subgraphs.
#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "jgpm_internal.hpp"
#include "plan.hpp"

struct jgpm_plan_s {
  jk::Plan plan{};
  int32_t order[jk::PLAN_MAXK] = {0};  // order[l] = pattern vertex matched at level l
  int64_t aut = 1;                     // |Aut(H)|
  std::vector<std::pair<int32_t, int32_t>> restrict_;  // (a, b): v_a < v_b, levels
};

namespace {

using Perm = std::vector<int>;

// all automorphisms of the pattern (k <= 7: at most 5040 permutations)
std::vector<Perm> automorphisms(int k, const uint32_t* adj) {
  std::vector<Perm> out;
  Perm p(k);
  std::iota(p.begin(), p.end(), 0);
  do {
    bool ok = true;
    for (int a = 0; a < k && ok; ++a)
      for (int b = 0; b < k && ok; ++b)
        if (((adj[a] >> b) & 1u) != ((adj[p[a]] >> p[b]) & 1u)) ok = false;
    if (ok) out.push_back(p);
  } while (std::next_permutation(p.begin(), p.end()));
  return out;
}

}  // namespace

extern "C" int jgpm_plan_compile(int32_t k, int32_t nedges, const int32_t* eu, const int32_t* ev, int induced,
                                 jgpm_plan_t* out) {
  if (!out || k < 2 || k > jk::PLAN_MAXK || nedges < 1 || !eu || !ev) return JGPM_E_INVALID;
  uint32_t adj[jk::PLAN_MAXK] = {0};
  for (int32_t e = 0; e < nedges; ++e) {
    const int a = eu[e], b = ev[e];
    if (a < 0 || a >= k || b < 0 || b >= k || a == b) return JGPM_E_INVALID;
    adj[a] |= 1u << b;
    adj[b] |= 1u << a;
  }
  auto deg = [&](int v) { return __builtin_popcount(adj[v]); };
  // matching order (also checks connectivity)
  int order[jk::PLAN_MAXK];
  uint32_t used = 0;
  int first = 0;
  for (int v = 1; v < k; ++v)
    if (deg(v) > deg(first)) first = v;
  order[0] = first;
  used = 1u << first;
  for (int l = 1; l < k; ++l) {
    int best = -1, bc = -1;
    for (int v = 0; v < k; ++v) {
      if ((used >> v) & 1u) continue;
      const int c = __builtin_popcount(adj[v] & used);
      if (c > bc || (c == bc && deg(v) > deg(best))) {
        best = v;
        bc = c;
      }
    }
    if (bc <= 0) return JGPM_E_PATTERN;  // not connected
    order[l] = best;
    used |= 1u << best;
  }
  // adjacency relabelled to levels
  int lev[jk::PLAN_MAXK];
  for (int l = 0; l < k; ++l) lev[order[l]] = l;
  uint32_t ladj[jk::PLAN_MAXK] = {0};
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b)
      if ((adj[a] >> b) & 1u) ladj[lev[a]] |= 1u << lev[b];
  auto* P = new (std::nothrow) jgpm_plan_s();
  if (!P) return JGPM_E_NOMEM;
  P->plan.k = k;
  P->plan.induced = induced ? 1 : 0;
  for (int l = 0; l < k; ++l) {
    P->order[l] = order[l];
    const uint32_t earlier = (1u << l) - 1;
    P->plan.conn[l] = ladj[l] & earlier;
    P->plan.nconn[l] = induced ? (~ladj[l] & earlier) : 0u;
  }
  // symmetry breaking on levels
  std::vector<Perm> G = automorphisms(k, ladj);
  P->aut = (int64_t)G.size();
  for (int l = 0; l < k && G.size() > 1; ++l) {
    std::vector<int> orbit;
    for (const Perm& g : G) orbit.push_back(g[l]);
    std::sort(orbit.begin(), orbit.end());
    orbit.erase(std::unique(orbit.begin(), orbit.end()), orbit.end());
    if (orbit.size() < 2) continue;
    for (int w : orbit)
      if (w != l) P->restrict_.push_back({l, w});  // v_l < v_w
    std::vector<Perm> stab;
    for (const Perm& g : G)
      if (g[l] == l) stab.push_back(g);
    G.swap(stab);
  }
  for (const auto& r : P->restrict_) {
    const int a = r.first, b = r.second;  // v_a < v_b, checked at the later level
    if (a < b) P->plan.gt[b] |= 1u << a;
    else P->plan.lt[a] |= 1u << b;
  }
  *out = P;
  return JGPM_OK;
}

extern "C" int jgpm_plan_describe(jgpm_plan_t p, int32_t* order, int64_t* aut_size, int32_t* ra, int32_t* rb,
                                  int32_t* nrestrict) {
  if (!p) return JGPM_E_INVALID;
  if (order) std::memcpy(order, p->order, sizeof(int32_t) * p->plan.k);
  if (aut_size) *aut_size = p->aut;
  if (nrestrict) {
    const int n = (int)p->restrict_.size();
    for (int i = 0; i < n && i < *nrestrict; ++i) {
      if (ra) ra[i] = p->restrict_[i].first;
      if (rb) rb[i] = p->restrict_[i].second;
    }
    *nrestrict = n;
  }
  return JGPM_OK;
}

extern "C" int jgpm_plan_free(jgpm_plan_t p) {
  if (!p) return JGPM_E_INVALID;
  delete p;
  return JGPM_OK;
}

namespace jgpm {
const jk::Plan* plan_of(const jgpm_plan_s* p) { return p ? &p->plan : nullptr; }
}  // namespace jgpm
