// Internal declarations shared by the host-side translation units of libjgpm.
#pragma once
#include <atomic>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

#include "jgpm.h"

// Host CSR: int64 row pointers and int32 ascending column indices, the
// on-disk and in-memory layout of the reference's Graph (include/graph.h:21-23).
struct jgpm_csr_s {
  int32_t nv = 0;
  int64_t nnz = 0;
  int32_t max_deg = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
};

namespace jgpm {

int host_threads();

// Dynamic-chunk parallel loop over [0, n) on host threads; f(begin, end).
void parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& f);

// splitmix64: the integer PRNG behind every generator (portable, seedable,
// counter-based so results do not depend on the thread count).  The same
// functions run in the device generators (ingest.hip), so a graph generated
// on the GPU is bit-identical to the host one.
#ifdef __HIP__
#define JGPM_HD __host__ __device__
#else
#define JGPM_HD
#endif
JGPM_HD static inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
JGPM_HD static inline uint64_t mix_seed(uint64_t seed, uint64_t stream) {
  uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull);
  return splitmix64(s);
}
// R-MAT quadrant thresholds: probabilities as 32-bit integer cut points.
struct RmatCuts {
  uint64_t ta, tab, tabc;
};
RmatCuts rmat_cuts(double a, double b, double c);
// One R-MAT sample (vertex bits drawn from sample i's own stream).
JGPM_HD static inline void rmat_sample(int scale, const RmatCuts& t, uint64_t seed, uint64_t i, uint32_t& u,
                                       uint32_t& v) {
  uint64_t st = mix_seed(seed, i);
  u = 0;
  v = 0;
  uint64_t bits = 0;
  for (int l = 0; l < scale; ++l) {
    if ((l & 1) == 0) bits = splitmix64(st);
    const uint64_t r = (l & 1) ? (bits >> 32) : (bits & 0xffffffffull);
    const uint32_t bu = r >= t.tab, bv = (r >= t.ta && r < t.tab) || r >= t.tabc;
    u = (u << 1) | bu;
    v = (v << 1) | bv;
  }
}

int build_sym_csr(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst, jgpm_csr_s& out);
// Checks what every kernel assumes of a CSR: row pointers from 0 to nnz,
// non-decreasing; ids in [0, nv); lists strictly ascending (sorted, no
// duplicates) without self loops; with `symmetric`, v in N(u) iff u in N(v).
// Returns JGPM_OK or `bad` (the caller's status for a malformed graph).
int check_csr(const jgpm_csr_s& c, bool symmetric, int bad);
int check_csr_arrays(int64_t nv, int64_t nnz, const int64_t* rp, const int32_t* col, bool symmetric, int bad);
// <prefix>.vertex.bin / .edge.bin mapped read-only (meta checked); unmap_csr
// releases them.
struct MappedCsr {
  int32_t nv = 0;
  int64_t nnz = 0;
  const int64_t* rp = nullptr;
  const int32_t* ci = nullptr;
  void* map_v = nullptr;
  void* map_e = nullptr;
  size_t len_v = 0, len_e = 0;
};
int map_csr(const char* prefix, MappedCsr& m);
void unmap_csr(MappedCsr& m);
void orient_dag(const jgpm_csr_s& sym, jgpm_csr_s& dag);
int32_t compute_max_degree(const jgpm_csr_s& c);
int write_csr(const jgpm_csr_s& c, const char* prefix);
int read_csr(const char* prefix, jgpm_csr_s& c);

}  // namespace jgpm
