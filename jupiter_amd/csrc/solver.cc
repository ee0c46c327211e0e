// SubgraphSolver — the one-GPU counting entry of the reference's driver
// (src/jupiter_gpm.cu:27, called from src/main.cc:46), declared in
// include/graph.h.  argv[2] names the pattern (P1..P16, TC, STAR3, TAILTRI,
// I_C4, I_DIAMOND, I_TAILTRI, I_STAR3, I_K4, or a pattern edge-list file: one
// isomorphic to a catalogue pattern runs its kernel, any other connected
// pattern of up to 7 vertices runs as a compiled plan); argv[3] (optional)
// is the number of rounds the root range is split into.  It prints the
// reference's two result lines (src/jupiter_gpm.cu:897,908) and, unlike the
// reference, stores the count in `total`.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "graph.h"
#include "jgpm.h"
#include "pattern.hh"

int jgpm_pattern_from_name(const std::string& s) {
  static const char* names[] = {"TC", "P1", "P2", "P3", "P4", "P5", "P6", "P7", "P8", "P9", "P10", "P11", "P12",
                                "P13", "P14", "P15", "P16", "STAR3", "TAILTRI"};
  for (int i = 0; i <= 18; ++i)
    if (s == names[i]) return i;
  if (s == "I_C4") return JGPM_I_C4;
  if (s == "I_DIAMOND") return JGPM_I_DIAMOND;
  if (s == "I_TAILTRI") return JGPM_I_TAILTRI;
  if (s == "I_STAR3") return JGPM_I_STAR3;
  if (s == "I_K4") return JGPM_I_K4;
  return -1;
}

int SubgraphSolver(Graph& g, uint64_t& total, int argc, char* argv[]) {
  std::string pname(argv[2]);
  int pid = jgpm_pattern_from_name(pname);
  jgpm_plan_t plan = nullptr;  // a pattern file outside the catalogue runs as a compiled plan
  if (pid < 0) {
    Pattern pat(pname, false);
    pid = pat.pattern_id();
    if (pid < 0) {
      std::map<vidType, int32_t> id;
      for (const auto& kv : pat.adjacency()) id.emplace(kv.first, (int32_t)id.size());
      std::vector<int32_t> eu, ev;
      for (const auto& kv : pat.adjacency())
        for (vidType u : kv.second)
          if (kv.first < u) {
            eu.push_back(id.at(kv.first));
            ev.push_back(id.at(u));
          }
      const int st = jgpm_plan_compile((int32_t)id.size(), (int32_t)eu.size(), eu.data(), ev.data(), 0, &plan);
      if (st) {
        std::cerr << "pattern " << pname << ": " << jgpm_strerror(st) << "\n";
        return st;
      }
      if (g.is_dag()) {
        jgpm_plan_free(plan);
        std::cerr << "pattern " << pname << " needs the symmetric graph\n";
        return JGPM_E_GRAPHKIND;
      }
    }
  }
  int total_round = argc >= 4 ? std::max(1, atoi(argv[3])) : 1;
  jgpm_ctx_t ctx = nullptr;
  int st = jgpm_init(0, &ctx);
  if (st) {
    jgpm_plan_free(plan);
    return st;
  }
  jgpm_graph_t dg = nullptr;
  st = jgpm_graph_upload(ctx, g.out_rowptr(), g.out_colidx(), g.num_vertices(), g.num_edges(),
                         g.is_dag() ? JGPM_DAG : JGPM_SYM, &dg);
  if (st) { jgpm_shutdown(ctx); return st; }
  // rounds split the roots into ranges of about equal edge count
  const eidType* rp = g.out_rowptr();
  const vidType nv = g.num_vertices();
  total = 0;
  double seconds = 0;
  vidType begin = 0;
  for (int r = 0; r < total_round && st == 0; ++r) {
    eidType target = (eidType)((double)g.num_edges() * (r + 1) / total_round);
    vidType end = (r == total_round - 1) ? nv : (vidType)(std::lower_bound(rp, rp + nv + 1, target) - rp);
    if (end < begin) end = begin;
    uint64_t c = 0;
    double ms = 0;
    st = plan ? jgpm_count_plan(dg, plan, begin, end, &c, &ms) : jgpm_count(dg, pid, begin, end, &c, &ms);
    total += c;
    seconds += ms * 1e-3;
    begin = end;
  }
  if (st == 0) {
    std::cout << "GPU:0 Kernel runtime = " << seconds << " sec << total_num_triangles = " << total << " \n";
    printf("=====================gpu count:%lu  runtime:%.6fs\n", (unsigned long)total, seconds);
  } else {
    std::cerr << "error: " << jgpm_strerror(st) << "\n";
  }
  jgpm_graph_free(dg);
  jgpm_shutdown(ctx);
  if (plan) jgpm_plan_free(plan);
  return st;
}

