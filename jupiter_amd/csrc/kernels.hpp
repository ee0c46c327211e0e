// Device kernel entry points of libjgpm (host-callable launchers).
#pragma once
#include "plan.hpp"
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jk {

// Root tiers of the clique engine, by DAG out-degree d of the root.
enum {
  TIER_WAVE = 0,   // d <= 64    one wave per root, one-word rows in LDS
  TIER_LDS2 = 1,   // d <= 128   block per root, rows in LDS
  TIER_LDS4 = 2,   // d <= 256
  TIER_LDS8 = 3,   // d <= 512
  TIER_LDS16 = 4,  // d <= 1024  (128 KB of rows in LDS, one block per CU)
  TIER_GLB32 = 5,  // d <= 2048  block per root, rows in a global scratch slice
  NUM_TIERS = 6
};

struct CliqueArgs {
  const int64_t* rp;     // DAG row pointers
  const int32_t* ci;     // DAG column indices
  const int32_t* ord;    // position of every vertex in a topological order of the DAG
  const int32_t* ids;    // task -> root vertex (NULL: root = rb + task)
  int64_t rb;
  int64_t ntasks;
  int32_t dmin, dmax;    // roots outside [dmin, dmax] are skipped by this launch
  int32_t ratio;         // a list longer than ratio * d is binary-searched, not streamed
  int32_t chunk;         // roots taken per queue grab (wave tier)
  uint32_t* qctr;        // work-queue counter (zero before launch)
  uint64_t* gscr;        // global row scratch (TIER_GLB32)
  uint64_t* out;         // uint64 accumulator
  uint32_t* rowctr;      // split 2048-wide count: one row counter per task, 8 queue
                         // counters before them (all zero before)
  int32_t rep;           // split 2048-wide count: blocks per task
};

// Query descriptor as the persistent kernel sees it.
struct BQ {
  const int64_t* rp;
  const int32_t* ci;
  const int32_t* ord;
  const int32_t* src;
  const int32_t* tri;
  const int64_t* nsum;
  int64_t base;    // first root (clique) or first edge id (P1..P12)
  int64_t ntasks;  // roots or edges
  int32_t kind;    // 3..7: clique size; 100 + P: symmetric pattern P
  int32_t chunk;   // roots / edges per chunk
};

hipError_t launch_clique(int k, bool bytes, int tier, const CliqueArgs& a, int grid, hipStream_t st);
// Split form of TIER_LDS16 for counting (k >= 4): build kernel (rows to
// a.gscr slots of split_slot_words() words per task), then count kernel.
hipError_t launch_clique_split(int k, bool build, const CliqueArgs& a, int grid, hipStream_t st);
size_t split_slot_words();
// Split form of TIER_GLB32 for k >= 5: build kernel (16 waves per root, rows
// to a.gscr slots of glb_slot_words() words per task), then count kernel
// (a.rep blocks per task share the root's rows through a.rowctr[task]).
hipError_t launch_clique_glb_split(int k, bool build, const CliqueArgs& a, int grid, hipStream_t st);
size_t glb_slot_words();

// Symmetric-graph patterns P1..P12: one wave per directed edge of [e0, e0+ne).
struct SymArgs {
  const int64_t* rp;     // symmetric row pointers
  const int32_t* ci;     // symmetric column indices
  const int32_t* src;    // COO source of every directed edge
  const int32_t* tri;    // per directed edge |N(u) ∩ N(v)| (patterns that use it)
  const int64_t* nsum;   // per vertex sum of neighbour degrees (P3)
  int64_t e0, ne;        // edge range
  unsigned long long* qctr;
  int32_t chunk;         // edges per queue grab
  int32_t* scr;          // per-wave list scratch for lists longer than the LDS capacity
  int64_t scr_stride;    // ints per scratch list (>= max degree)
  uint64_t* out;
  // P5, P7, P9, P11, P12: edges with heavy_min < tri[e] <= sym_heavy_nmax()
  // go to launch_sym_heavy (0: every edge runs here; needs tri)
  int32_t heavy_min;
  int32_t heavy_stream;  // stream N(v2) when d(v2) <= heavy_stream * |S|, else binary searches
  int32_t nv;            // vertices (the heavy kernel keeps S as a bitmap on small graphs)
  // the wave-per-edge launch appends each heavy edge to list 0 (|S| <=
  // heavy_mid), 1 (|S| <= sym_heavy_nmax()) or 2 (beyond, on graphs of at
  // most 2^20 vertices), at hlist + list * hcap, counters hctr[0..2]; an edge
  // that finds its list full it counts itself.  The heavy launches pop with
  // hctr[3..5].
  int32_t* hlist;
  unsigned long long* hctr;
  int32_t hcap;
  int32_t heavy_mid;
};
hipError_t launch_sym(int p, const SymArgs& a, int grid, hipStream_t st);
// Whole-graph P3 from one intersection per undirected edge (each triangle
// once, its three oriented terms); needs tri and nsum, e0 = 0, ne = all edges.
hipError_t launch_p3_triangles(const SymArgs& a, int grid, hipStream_t st);
// The heavy edges of P5, P7, P9, P11 and P12, one block per edge, popped
// from list `shape` (0, 1, 2 as above) that launch_sym filled; shape 2 keeps
// S in the block's slice of a.scr (a.scr_stride ints, grid slices).
bool sym_heavy_pattern(int p);
int sym_heavy_nmax();
int sym_heavy_stream();
int sym_heavy_mid(int p);  // |S| up to this: the narrow launch (more blocks per CU)
hipError_t launch_sym_heavy(int p, const SymArgs& a, int shape, int grid, hipStream_t st);

// 4-cycles (P1) by wedges: one wave per root of [rb, re), counters in a
// per-wave dense array of nv ints (cnt + wave * nv), zero before and after.
struct P1Args {
  const int64_t* rp;
  const int32_t* ci;
  int64_t rb, re;
  int32_t nv;
  int32_t desc;  // deal the roots from re - 1 down (degree-ordered graphs: widest roots first)
  unsigned* qctr;
  int32_t* cnt;
  uint64_t* out;
};
constexpr int P1_WPB = 4;
hipError_t launch_p1_wedge(const P1Args& a, int grid, hipStream_t st);
// Every 4-cycle query of a batch (BQ kind 101: rp/ci, first root `base`) in
// one launch over the flattened roots; p1cs[i] = roots of the 4-cycle
// queries before query i (nq + 1 entries), counts[i] the query's slot.
struct P1BatchArgs {
  const BQ* q;
  const int64_t* p1cs;
  int nq;
  int64_t total;
  unsigned* qctr;        // zero before the launch
  int32_t* cnt;          // per wave cnt_stride ints, zero before and after
  int64_t cnt_stride;    // >= the largest vertex count of the queries
  uint64_t* counts;
};
hipError_t launch_p1_batch(const P1BatchArgs& a, int grid, hipStream_t st);

// One wave per root vertex v_0 of [rb, re) on the symmetric CSR.
struct PlanArgs {
  const int64_t* rp;
  const int32_t* ci;
  int64_t rb, re;
  unsigned long long* qctr;
  int32_t* scr;                  // per wave (k + 1) lists of `stride` ints
  int64_t stride;                // >= maximum degree
  uint64_t* out;
  Plan plan;
};
constexpr int PLAN_WPB = 4;
hipError_t launch_plan(const PlanArgs& a, int grid, hipStream_t st);
int sym_waves_per_block();
int sym_lists_per_wave();
hipError_t launch_aux_src(const int64_t* rp, int32_t nv, int32_t* src, int grid, hipStream_t st);
hipError_t launch_aux_tri(const int64_t* rp, const int32_t* ci, const int32_t* src, int64_t ne, int32_t* tri,
                          unsigned long long* qctr, int grid, hipStream_t st);
hipError_t launch_aux_nsum(const int64_t* rp, const int32_t* ci, int32_t nv, int64_t* nsum, int grid, hipStream_t st);
// mode 0: sum C(d,3) (3-stars); mode 1: sum t_v (d_v - 2) (tailed triangles)
hipError_t launch_vertex_sums(const int64_t* rp, const int32_t* tri, int64_t rb, int64_t re, int mode,
                              unsigned long long* out, int grid, hipStream_t st);
// Diamonds over the directed edges [e0, e1) from the per-edge triangle array.
hipError_t launch_p2_edges(const int32_t* src, const int32_t* ci, const int32_t* tri, int64_t e0, int64_t e1,
                           unsigned long long* out, int grid, hipStream_t st);
int tier_max_deg(int tier);
size_t tier_scratch_words(int tier);

// Roots wider than the widest row tier (DAG out-degree > 2048): the level
// nest per root, lists in global scratch (clique_huge.hip).
struct HugeArgs {
  const int64_t* rp;     // DAG row pointers
  const int32_t* ci;     // DAG column indices
  const int32_t* ids;    // the roots
  int64_t ntasks;
  int32_t* scr;          // huge_lists_per_block(k) * grid lists of stride ints
  int64_t stride;        // >= the DAG's maximum out-degree
  uint64_t* out;
};
int huge_lists_per_block(int k);
hipError_t launch_clique_huge(int k, const HugeArgs& a, int grid, hipStream_t st);

}  // namespace jk
