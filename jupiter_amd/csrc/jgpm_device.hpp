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
  unsigned long long* d_q64 = nullptr;  // CTX_SLOTS 64-bit work-queue counters
  int32_t* d_sscr = nullptr;  // list scratch of the symmetric-pattern kernels
  size_t sscr_ints = 0;
  int32_t* d_hlist = nullptr;  // heavy-edge lists of P5/P7/P9/P11/P12 (3 x hcap ints)
  int64_t hcap = 0;
  unsigned long long* d_hctr = nullptr;  // their append and pop counters (6)
  int32_t* d_p1cnt = nullptr;  // wedge counters of the 4-cycle kernel (all zero between launches)
  size_t p1cnt_ints = 0;
  // batched / served launches: query table, chunk prefix, per-query counts
  void* d_bq = nullptr;
  int64_t* d_bcs = nullptr;
  uint64_t* d_bcnt = nullptr;
  unsigned* d_bq32 = nullptr;  // queue counters of the follow-up tier launches
  int bcap = 0;
  int batch_occ = 0;  // resident blocks per CU of the batched kernel (occupancy)
  std::vector<jk::BQ> bq_host; // staging of the last batched launch
  std::vector<int64_t> bcs_host;
  uint64_t* d_scr = nullptr;  // row scratch of the global-row tier
  size_t scr_words = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // per-launch profile of the current query (events around every kernel)
  static constexpr int MAX_PROF = 16;
  hipEvent_t pev[2 * MAX_PROF] = {nullptr};
  const char* pname[MAX_PROF] = {nullptr};
  int nprof = 0;
  bool prof_on = true;
  bool capturing = false;  // launches are being captured into a HIP graph: nothing may allocate
  std::mutex mu;
};

struct jgpm_graph_s {
  jgpm_ctx_s* ctx = nullptr;
  bool has_sym = false, has_dag = false;
  DevCSR sym, dag;
  std::vector<int32_t> tier_h[jk::NUM_TIERS];  // root ids per block tier (ascending)
  int32_t* tier_d[jk::NUM_TIERS] = {nullptr};
  int32_t* d_ord = nullptr;                    // topological position of every vertex in the DAG
  std::vector<int32_t> huge_h;                 // roots beyond the largest tier (ascending)
  int32_t* huge_d = nullptr;                   // their device copy (first count that needs it)
  std::vector<int64_t> sym_rp_h;               // host copy of the symmetric row pointers
  std::vector<int64_t> dag_rp_h;               // host copy of the DAG row pointers
  int32_t sym_mmin = 1;                        // max over edges of min(d(u), d(v)): longest S = N(u) ∩ N(v)
  int32_t* d_src = nullptr;                    // COO source per symmetric edge (lazy)
  int32_t* d_tri = nullptr;                    // triangles per symmetric edge (lazy)
  int64_t* d_nsum = nullptr;                   // neighbour-degree sums (lazy)
  DevCSR rsym;                                 // symmetric CSR relabelled by d_ord (lazy, 4-cycle totals)
  int64_t bytes = 0;
  int nbatch = 0;                              // prepared batches that reference this graph
};

namespace jgpm {
// The context whose stream, accumulators, counters and scratch the dispatch
// routines use for graph g: g's own context, unless an ExecScope on the
// calling thread names another context of the same device (the serving
// loop's heavy-request lanes run graphs resident on the server's context on
// lane contexts of their own, so their launches overlap on the device).
jgpm_ctx_s* exec_ctx(const jgpm_graph_s* g);
struct ExecScope {
  jgpm_ctx_s* prev;
  explicit ExecScope(jgpm_ctx_s* c);
  ~ExecScope();
  ExecScope(const ExecScope&) = delete;
  ExecScope& operator=(const ExecScope&) = delete;
};
// Builds every per-graph array that `pattern` over roots [rb, re) would
// otherwise build on first use (COO sources, per-edge triangles, neighbour
// degree sums, the degree-order relabelling, the wide-root ids), on g's own
// context under its lock, and waits for them: after it, a count of that
// query allocates and writes nothing of g's, so it may run on another
// context's stream while other work on g proceeds.
int prepare_pattern(jgpm_graph_s* g, int pattern, int64_t rb, int64_t re);
// Device copy of the ids of the roots wider than the widest row tier
// (g->huge_d), made on first use.
int ensure_huge_ids(jgpm_graph_s* g);
int ctx_reset_counters(jgpm_ctx_s* c, hipStream_t st);
// p[0..n) = 0 on st, by a fill kernel (no memset node: callers may be captured).
int zero_u64(uint64_t* p, int n, hipStream_t st);
// Address of this device's static scratch array of jgpm_count_noalloc
// (noalloc.hip), taken once per device with `dev` current.
int noalloc_setup(int dev, uint64_t** out);
// 4-cycle (P1) wedge kernel.  p1_counter_ints: the zeroed counter ints a
// graph of nv vertices needs; ensure_p1_counters grows the context's (it
// allocates, never while capturing); launch_p1 only launches, into *out, with
// the zeroed 32-bit queue counter qctr and counters cnt[0..cnt_ints) that are
// all zero (and are left so).
size_t p1_counter_ints(jgpm_ctx_s* c, int32_t nv);
// g->rsym: vertex v becomes d_ord[v] (degree order), lists re-sorted.  Built
// once per graph (allocates; never while capturing).
int ensure_rank_csr(jgpm_graph_s* g, hipStream_t st);
// A checked symmetric CSR in host memory -> resident graph; the degree DAG is
// built on the device (ingest.hip).  dag_only keeps only the DAG.
int graph_from_host_sym(jgpm_ctx_s* c, const int64_t* rowptr, const int32_t* colidx, int32_t nv, int64_t nnz,
                        bool dag_only, jgpm_graph_t* out);
int ensure_p1_counters(jgpm_ctx_s* c, size_t need, hipStream_t st);
int launch_p1(jgpm_graph_s* g, int64_t rb, int64_t re, uint64_t* out, unsigned* qctr, int32_t* cnt, size_t cnt_ints,
              hipStream_t st);
int ensure_scratch(jgpm_ctx_s* c, size_t words);
int clique_k(int pattern);
int enqueue_clique(jgpm_graph_s* g, int k, bool bytes, int64_t rb, int64_t re, int out_slot, hipStream_t st,
                   bool split_slots = false, int first_tier = 0, uint64_t* out_ptr = nullptr,
                   unsigned* qctr_base = nullptr, uint64_t* gscr = nullptr, size_t gscr_words = 0);
// Blocks of the global-row tier one launch may use, and their row scratch.
int glb_tier_grid(const jgpm_ctx_s* c);
// Root tier lists of the DAG (host row pointers) and the degree order.
int build_tiers(jgpm_graph_s* g, const int64_t* rp, int32_t nv);
int set_degree_order(jgpm_graph_s* g, const std::vector<int64_t>& sym_rp, int32_t max_deg);
int ensure_sym_aux(jgpm_graph_s* g, bool need_tri, bool need_nsum, hipStream_t st);
// Batched persistent launch over queries the kernel supports (TC, P1..P16);
// counts land in c->d_bcnt[0..n).  hq / cs are host staging that must stay
// alive until the stream is synchronised.
int enqueue_batch_q(jgpm_ctx_s* c, const jgpm_query* q, int n, hipStream_t st);
bool batch_supported(int pattern);
bool pattern_uses_tri(int p);
// whether a single-query launch of pattern p on context c reads the triangle array
bool query_needs_tri(const jgpm_ctx_s* c, int p);
// profile hooks around one kernel launch on the context stream
void prof_begin(jgpm_ctx_s* c, hipStream_t st, const char* name);
void prof_end(jgpm_ctx_s* c, hipStream_t st);
int enqueue_sym(jgpm_graph_s* g, int p, int64_t rb, int64_t re, int out_slot, hipStream_t st);
int enqueue_vertex_sum(jgpm_graph_s* g, int mode, int64_t rb, int64_t re, int out_slot, hipStream_t st);
// Enqueue every kernel of `pattern` for roots [rb, re); results land in
// accumulator slots base_slot.. (see finish_pattern).
int enqueue_pattern(jgpm_graph_s* g, int pattern, int64_t rb, int64_t re, int base_slot, hipStream_t st);
// Number of accumulator slots a pattern uses.
int pattern_slots(int pattern);
// Combine the accumulator slots of a finished query into its count.
uint64_t finish_pattern(int pattern, const uint64_t* slots);
}  // namespace jgpm
