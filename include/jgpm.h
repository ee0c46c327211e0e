/*
 * jgpm.h — C ABI of the MI355X (gfx950) graph pattern counting library.
 *
 * Plain C types only: pointers, sizes and integer status codes.  Every entry
 * point returns JGPM_OK (0) or a negative JGPM_E_* status; nothing in the
 * library calls exit() or abort().  jgpm_strerror() maps a status to text.
 *
 * What each entry point replaces in the reference (AnySparse/Jupiter):
 *   jgpm_init / jgpm_shutdown      cudaSetDevice + nvshmemx_init_attr in
 *                                  SubgraphSolver (src/jupiter_gpm.cu:39-58)
 *   jgpm_graph_upload / _load      Graph(prefix, use_dag) (common/graph.cc:8-124)
 *                                  + NvsGraphGPU::init (include/graph_gpu_nvs.h:258-304)
 *   jgpm_count                     the round loop + name->kernel dispatch + count
 *                                  readback (src/jupiter_gpm.cu:356-908)
 *   jgpm_count_batch, jgpm_batch_*,
 *   jgpm_autotune_batch            no reference counterpart (one query per run there;
 *                                  the reference's host chunk scheduler,
 *                                  common/scheduler.cc, is never called)
 *   jgpm_serve_*                   no reference counterpart (continuous batching)
 *   jgpm_plan_*, jgpm_count_plan   the general pattern generator the reference leaves as
 *                                  a TODO (src/gpu_kernels_codegen.cuh:3-4); levels as
 *                                  in PatternConfig (src/config.h:39-58)
 *   jgpm_graph_from_edges /
 *   jgpm_graph_rmat                graph construction + orientation on the GPU
 *                                  (common/graph.cc:8-124,210-401 do it on the host)
 *   jgpm_gen_* / jgpm_csr_*        host graph construction used by the CLI, tests
 *                                  and benchmarks; orientation follows
 *                                  Graph::orientation (common/graph.cc:210-401)
 */
#ifndef JGPM_H
#define JGPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define JGPM_OK              0
#define JGPM_E_INVALID      -1   /* bad argument (null pointer, bad range, bad flags) */
#define JGPM_E_PATTERN      -2   /* unknown pattern id (reference: assert(0), src/jupiter_gpm.cu:428,872) */
#define JGPM_E_NOMEM        -3   /* device or host allocation failed */
#define JGPM_E_HIP          -4   /* a HIP runtime call failed */
#define JGPM_E_IO           -5   /* file missing / short read / bad meta file */
#define JGPM_E_BUDGET       -6   /* request exceeds the HBM budget of the serving loop */
#define JGPM_E_GRAPHKIND    -7   /* pattern needs the symmetric CSR but the graph was uploaded as DAG-only */
#define JGPM_E_BUSY         -8   /* serving queue full (retry), or a graph still used by a prepared batch */
#define JGPM_E_NOTREADY     -9   /* serving result not ready yet */
#define JGPM_E_STATE       -10   /* call not valid in the current state */
#define JGPM_E_NODEVICE    -11   /* no HIP device available */

/* ---- pattern ids ------------------------------------------------------- *
 * P1..P16 keep the reference's names and meanings (the kernels of
 * src/message_passing_gpu_kernels.cuh:104-1537).  Each equals the
 * non-induced subgraph count of a fixed pattern graph H:
 *   P1 4-cycle, P2 diamond, P3 triangle with a 2-path tail, P4 house,
 *   P5 K4 plus a vertex on an edge, P6 K_{2,3} plus an edge, P7 gem,
 *   P8 wheel W4, P9 K5 minus an edge, P10 5-fan, P11 3-sun,
 *   P12 the (6 vertex, 11 edge) graph of P12_kernel, P13..P16 K4..K7.
 * Extension ids (no reference P-id):
 *   JGPM_TC        triangles (reference P0_kernel / total[1])
 *   JGPM_STAR3     non-induced 3-stars, sum_v C(d_v,3)
 *   JGPM_TAILTRI   non-induced tailed triangles, sum_v t_v (d_v - 2)
 *   JGPM_I_*       vertex-induced 4-vertex motifs obtained from the
 *                  non-induced counts by the exact inverse conversion matrix
 */
enum {
  JGPM_TC = 0,
  JGPM_P1 = 1, JGPM_P2, JGPM_P3, JGPM_P4, JGPM_P5, JGPM_P6, JGPM_P7, JGPM_P8,
  JGPM_P9, JGPM_P10, JGPM_P11, JGPM_P12, JGPM_P13, JGPM_P14, JGPM_P15, JGPM_P16,
  JGPM_STAR3 = 17,
  JGPM_TAILTRI = 18,
  JGPM_I_C4 = 20,      /* induced 4-cycle            */
  JGPM_I_DIAMOND = 21, /* induced diamond            */
  JGPM_I_TAILTRI = 22, /* induced tailed triangle    */
  JGPM_I_STAR3 = 23,   /* induced 3-star             */
  JGPM_I_K4 = 24,      /* induced K4 (== P13)        */
  JGPM_NUM_PATTERNS = 25
};

/* graph upload flags */
#define JGPM_SYM 1u   /* the CSR is symmetric (undirected); the library derives the degree DAG */
#define JGPM_DAG 2u   /* the CSR is already the degree-oriented DAG; only DAG patterns allowed */

typedef struct jgpm_ctx_s*   jgpm_ctx_t;
typedef struct jgpm_graph_s* jgpm_graph_t;
typedef struct jgpm_csr_s*   jgpm_csr_t;
typedef struct jgpm_serve_s* jgpm_serve_t;

/* ---- context ----------------------------------------------------------- */
/* Opens HIP device `dev`, creates the context's stream and work buffers. */
int  jgpm_init(int dev, jgpm_ctx_t* out);
int  jgpm_shutdown(jgpm_ctx_t ctx);
const char* jgpm_strerror(int status);
/* Device properties: CU count, LDS bytes per workgroup, HBM bytes. */
int  jgpm_device_info(jgpm_ctx_t ctx, int* num_cu, int* lds_bytes, int64_t* hbm_bytes,
                      char* name, int name_len);

/* ---- graphs ------------------------------------------------------------ */
/* Copies a host CSR (int64 rowptr[nv+1], int32 colidx[nnz], ascending lists)
 * to the device.  The caller keeps ownership of the host arrays. */
int  jgpm_graph_upload(jgpm_ctx_t ctx, const int64_t* rowptr, const int32_t* colidx,
                       int32_t nv, int64_t nnz, uint32_t flags, jgpm_graph_t* out);
/* Loads <prefix>.meta.txt / .vertex.bin / .edge.bin (format of
 * Graph::Graph and Graph::dump_binary, common/graph.cc:26-46,727-753).
 * dag != 0 keeps only the degree DAG (as Graph(prefix, use_dag=true)). */
int  jgpm_graph_load(jgpm_ctx_t ctx, const char* prefix, int dag, jgpm_graph_t* out);
int  jgpm_graph_free(jgpm_graph_t g);
/* Sizes of the resident graph: vertices, symmetric nnz, DAG nnz, max degrees. */
int  jgpm_graph_info(jgpm_graph_t g, int32_t* nv, int64_t* sym_nnz, int64_t* dag_nnz,
                     int32_t* sym_max_deg, int32_t* dag_max_deg);
/* Device bytes held by the graph handle. */
int  jgpm_graph_bytes(jgpm_graph_t g, int64_t* bytes);

/* ---- counting ---------------------------------------------------------- */
/* Count pattern `pattern` over the roots [root_begin, root_end) — the
 * source vertices v0 of the matched edges, as the reference's per-GPU edge
 * ranges (src/jupiter_gpm.cu:153-162,373-375).  root_end = -1 means nv.
 * Blocks until the count is on the host.  kernel_ms (may be NULL) receives
 * the device time between the first and last kernel of the query.
 * JGPM_I_C4, JGPM_I_DIAMOND, JGPM_I_TAILTRI and JGPM_I_STAR3 combine counts
 * that credit an occurrence to different vertices, so they are defined on the
 * whole graph only: any other range returns JGPM_E_INVALID. */
int  jgpm_count(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end,
                uint64_t* count, double* kernel_ms);
/* jgpm_count for the clique patterns (JGPM_TC, P13..P16, JGPM_I_K4) without
 * any device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The query's device scratch (accumulators,
 * queue counters, rows of the 2048-wide tier) is a static 32 MiB array per
 * GPU that the library holds until the process ends; the 2048-wide tier runs
 * with as many blocks as it holds row slices, and roots wider than 2048 keep
 * their level lists there too.  Other pattern ids return JGPM_E_PATTERN.  The
 * array is shared by every call on one GPU: calls through one context are
 * serialised, calls through two contexts on the same GPU must not overlap. */
int  jgpm_count_noalloc(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end,
                        uint64_t* count);
/* Algorithmic bytes of the reference's canonical loop nest for the same
 * query: every neighbour-list fetch costs 16 + 4*d(v) bytes.  Computed on
 * the device.  Defined for JGPM_TC and P13..P16. */
int  jgpm_work_bytes(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end,
                     uint64_t* bytes);

/* Per-kernel profile of the last jgpm_count on this context: the device
 * time (HIP events on the context stream) of every kernel launch, in launch
 * order, with a static name for each.  Returns the number of launches. */
int  jgpm_last_profile(jgpm_ctx_t ctx, int n, const char** names, double* ms);
/* jgpm_work_bytes split over the same kernel launches as jgpm_count uses
 * for the query (order matches jgpm_last_profile).  Returns the number. */
int  jgpm_work_bytes_split(jgpm_graph_t g, int pattern, int64_t root_begin, int64_t root_end,
                           uint64_t* bytes, int n);

/* Tuning knobs (autotuner output).  In jgpm_set_tuning a field > 0 sets the
 * knob (clamped to its range), 0 keeps the current value, < 0 restores the
 * default. */
typedef struct {
  /* small_max_deg, lds_max_deg and waves_per_block are kept and reported but
   * not read by the kernels: the clique tiers are compiled for out-degrees
   * 64 (one wave per root, 4 roots per block), 128, 256, 512, 1024 (bitmap
   * rows in LDS) and 2048 (rows in global scratch). */
  int32_t small_max_deg;    /* (fixed: 64) roots that use one wave each                  */
  int32_t lds_max_deg;      /* (fixed: 1024) roots that keep their bitmap in LDS          */
  int32_t gallop_ratio;     /* neighbour list longer than ratio*d(root) -> binary search
                               (ratio/4 for roots of out-degree <= 128, ratio/8 <= 256)     */
  int32_t waves_per_block;  /* (fixed: 4) waves per workgroup of the wave-per-root kernel  */
  int32_t blocks_per_cu;    /* persistent grid size = blocks_per_cu * num_cu              */
  /* batched launches (jgpm_count_batch, jgpm_batch_*, serving) */
  int32_t batch_hub_deg;    /* clique roots above this out-degree leave the wave-per-root
                               persistent kernel for block-per-hub launches: 64 (one-word
                               LDS rows) or 128 (two-word rows)                             */
  int32_t batch_root_chunk; /* roots per work item of a clique query                      */
  int32_t batch_edge_chunk; /* directed edges per work item of a P1..P12 query (P2 and P3,
                               whose per-edge work is O(1) after the intersection, use 8x) */
  int32_t batch_blocks_per_cu; /* persistent batched grid = this * num_cu, capped by the
                                  kernel's occupancy                                       */
  int32_t sym_heavy_min;    /* P5, P7, P9, P11, P12 (single-query launches): directed edges
                               with more than this many common neighbours (up to 8192) run
                               one block per edge, S in an LDS hash (default 32; 8192
                               turns it off)                                               */
} jgpm_tuning;
int  jgpm_get_tuning(jgpm_ctx_t ctx, jgpm_tuning* t);
int  jgpm_set_tuning(jgpm_ctx_t ctx, const jgpm_tuning* t);

/* ---- on-device ingest ---------------------------------------------------- */
/* Undirected edge list (host arrays of m endpoint pairs, ids in [0, nv)) built
 * into a resident graph on the GPU: self loops dropped, both directions,
 * ascending duplicate-free lists (device radix sort), and the degree DAG
 * (rule of common/graph.cc:229-230).  The same graph as jgpm_csr_from_edges
 * followed by jgpm_graph_upload(JGPM_SYM).  Requires m < 2^30. */
int  jgpm_graph_from_edges(jgpm_ctx_t ctx, int32_t nv, int64_t m, const int32_t* src,
                           const int32_t* dst, jgpm_graph_t* out);
/* Seeded R-MAT sampled and built on the GPU: the same graph as jgpm_gen_rmat
 * (identical samples) uploaded with JGPM_SYM.  Requires edge_factor << scale
 * < 2^30. */
int  jgpm_graph_rmat(jgpm_ctx_t ctx, int scale, int edge_factor, double a, double b, double c,
                     uint64_t seed, jgpm_graph_t* out);
/* Copy a resident graph's symmetric (dag = 0) or DAG (dag = 1) CSR back to a
 * host CSR handle (free with jgpm_csr_free; jgpm_csr_write saves it in the
 * reference's file format). */
int  jgpm_graph_download(jgpm_graph_t g, int dag, jgpm_csr_t* out);

/* ---- general patterns: pattern graph -> compiled plan -> count ---------- */
/* Compile a connected pattern with k (2..7) vertices and the given edges into
 * a matching plan: matching order, adjacency (and, with induced != 0,
 * non-adjacency) constraints per level, and symmetry-breaking restrictions
 * from the pattern's automorphism group, so that every occurrence is matched
 * once.  induced = 0 counts subgraphs (the P-id semantics), 1 counts
 * vertex-induced subgraphs.  JGPM_E_PATTERN if the pattern is not connected. */
typedef struct jgpm_plan_s* jgpm_plan_t;
int  jgpm_plan_compile(int32_t k, int32_t nedges, const int32_t* eu, const int32_t* ev, int induced,
                       jgpm_plan_t* out);
/* order[l] (k entries): pattern vertex matched at level l; aut_size = |Aut|;
 * restrictions (ra[i], rb[i]) mean v_ra < v_rb over levels.  *nrestrict is
 * the capacity on entry and the count on return.  Any pointer may be NULL. */
int  jgpm_plan_describe(jgpm_plan_t p, int32_t* order, int64_t* aut_size, int32_t* ra, int32_t* rb,
                        int32_t* nrestrict);
/* Count occurrences on the symmetric CSR; a root range selects the vertices
 * matched at level 0 (so disjoint ranges add up to the total). */
int  jgpm_count_plan(jgpm_graph_t g, jgpm_plan_t p, int64_t root_begin, int64_t root_end,
                     uint64_t* count, double* kernel_ms);
int  jgpm_plan_free(jgpm_plan_t p);

/* ---- batching: many (graph, pattern, range) queries in one launch ------ */
typedef struct {
  jgpm_graph_t graph;
  int32_t pattern;
  int32_t _pad;
  int64_t root_begin;
  int64_t root_end;      /* -1 = nv */
} jgpm_query;
int  jgpm_count_batch(jgpm_ctx_t ctx, const jgpm_query* queries, int n, uint64_t* counts,
                      double* kernel_ms);

/* Prepared batch: the query table, its buffers and every launch of one batched
 * count (counter fills, the persistent kernel, the block-per-hub launches)
 * captured once in a HIP graph and replayed by jgpm_batch_run.  The tuning in
 * force at create time is baked in.  Queries must be TC or P1..P16.  A graph
 * referenced by a live batch cannot be freed (JGPM_E_BUSY); free the batch
 * first.  Runs on the context stream; one thread per context at a time. */
typedef struct jgpm_batch_s* jgpm_batch_t;
int  jgpm_batch_create(jgpm_ctx_t ctx, const jgpm_query* queries, int n, jgpm_batch_t* out);
int  jgpm_batch_run(jgpm_batch_t b, uint64_t* counts, double* kernel_ms);
int  jgpm_batch_free(jgpm_batch_t b);

/* Autotuner for batched launches: times the batch (device time, HIP events,
 * best of `reps`) under candidate settings of batch_hub_deg, batch_root_chunk,
 * batch_edge_chunk and batch_blocks_per_cu (coordinate descent from the current
 * tuning), requires every candidate's counts to equal the first run's, and
 * installs the fastest setting in the context.  best / best_ms may be NULL.
 * Returns JGPM_E_STATE if two candidates disagree on a count. */
int  jgpm_autotune_batch(jgpm_ctx_t ctx, const jgpm_query* queries, int n, int reps,
                         jgpm_tuning* best, double* best_ms);

/* ---- serving: continuous batching over device-resident graphs ---------- */
typedef struct {
  int64_t hbm_budget_bytes;  /* admission budget for per-query intermediates   */
  int32_t max_inflight;      /* queries admitted to the device at once         */
  int32_t slice_roots;       /* roots per scheduling slice (fairness)          */
  int32_t heavy_lanes;       /* heavy queries running at once, each on a stream
                                and scratch of its own (0: 4)                  */
  int32_t reserved;          /* zero                                           */
} jgpm_serve_config;
int  jgpm_serve_start(jgpm_ctx_t ctx, const jgpm_serve_config* cfg, jgpm_serve_t* out);
/* Enqueue a query; returns a ticket.  Thread-safe. */
int  jgpm_serve_submit(jgpm_serve_t s, jgpm_graph_t g, int pattern, int64_t root_begin,
                       int64_t root_end, uint64_t* ticket);
/* Non-blocking: JGPM_E_NOTREADY until the ticket's count is available.
 * latency_us (may be NULL): submit -> result-visible time.  A result is
 * handed out once (poll or wait); the server then forgets the ticket, and a
 * second retrieval returns JGPM_E_INVALID. */
int  jgpm_serve_poll(jgpm_serve_t s, uint64_t ticket, uint64_t* count, double* latency_us);
int  jgpm_serve_wait(jgpm_serve_t s, uint64_t ticket, uint64_t* count, double* latency_us);
int  jgpm_serve_stop(jgpm_serve_t s);

/* ---- host CSR construction (generators, orientation, files) ------------ */
/* Seeded R-MAT: 2^scale vertices, edge_factor * 2^scale samples, quadrant
 * probabilities a, b, c (d = 1-a-b-c); self loops dropped, symmetrised,
 * deduplicated, ascending lists.  Bit-reproducible for a given seed. */
int  jgpm_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                   jgpm_csr_t* out);
/* Chung-Lu power law: weights (i + i0)^(-1/(gamma-1)) scaled to target
 * avg edges and max weight; alias-method endpoint sampling. */
int  jgpm_gen_chung_lu(int32_t n, int64_t target_edges, double gamma, double max_weight,
                       uint64_t seed, jgpm_csr_t* out);
/* Erdos-Renyi with the given mean degree (uniform pair sampling). */
int  jgpm_gen_er(int32_t n, double mean_degree, uint64_t seed, jgpm_csr_t* out);
/* Barabasi-Albert preferential attachment, m edges per new vertex. */
int  jgpm_gen_ba(int32_t n, int32_t m, uint64_t seed, jgpm_csr_t* out);
/* Build a symmetric CSR from an arbitrary undirected edge list. */
int  jgpm_csr_from_edges(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                         jgpm_csr_t* out);
/* Degree-ordered DAG of a symmetric CSR: keep u->v iff d(v) > d(u) or
 * (d(v) == d(u) and v > u) (common/graph.cc:229-230). */
int  jgpm_csr_orient(jgpm_csr_t sym, jgpm_csr_t* dag);
int  jgpm_csr_info(jgpm_csr_t c, int32_t* nv, int64_t* nnz, int32_t* max_deg);
int  jgpm_csr_data(jgpm_csr_t c, const int64_t** rowptr, const int32_t** colidx);
int  jgpm_csr_write(jgpm_csr_t c, const char* prefix);
int  jgpm_csr_read(const char* prefix, jgpm_csr_t* out);
int  jgpm_csr_free(jgpm_csr_t c);

#ifdef __cplusplus
}
#endif
#endif /* JGPM_H */
