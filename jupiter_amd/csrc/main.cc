// jupiter_gpm — drop-in command line of the reference (src/main.cc:7-48):
//   jupiter_gpm <graph_prefix> <pattern> [total_round=1]
// <pattern> is P1..P16 (or TC, STAR3, TAILTRI, I_C4, I_DIAMOND, I_TAILTRI,
// I_STAR3, I_K4, or a pattern edge-list file: one isomorphic to a catalogue
// pattern runs its kernel, any other connected pattern of up to 7 vertices
// runs as a compiled plan).  <graph_prefix> names the
// .meta.txt/.vertex.bin/.edge.bin files; "rmat:<scale>[:<seed>]" generates a
// seeded R-MAT graph (edge factor 16, a/b/c = 0.57/0.19/0.19) instead.
// P13..P16 load the degree DAG, as in the reference.  The final lines match
// the reference's output (src/jupiter_gpm.cu:897,908).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "graph.h"
#include "jgpm.h"
#include "pattern.hh"

static int pattern_from_name(const std::string& s) {
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

static bool dag_pattern(int p) { return p == JGPM_TC || (p >= JGPM_P13 && p <= JGPM_P16) || p == JGPM_I_K4; }

int SubgraphSolver(Graph& g, uint64_t& total, int argc, char* argv[]) {
  std::string pname(argv[2]);
  int pid = pattern_from_name(pname);
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

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cout << "Usage: " << argv[0] << " <graph> [pattern(P1)] [round(1)]\n";
    std::cout << "Example: " << argv[0] << " /graph_inputs/mico/graph P1\n";
    return 1;
  }
  std::string pname(argv[2]);
  int pid = pattern_from_name(pname);
  const bool use_dag = pid >= JGPM_P13 && pid <= JGPM_P16;
  std::string src(argv[1]);
  try {
    std::unique_ptr<Graph> gp;
    if (src.rfind("rmat:", 0) == 0) {
      int scale = atoi(src.c_str() + 5);
      const char* s2 = strchr(src.c_str() + 5, ':');
      uint64_t seed = s2 ? strtoull(s2 + 1, nullptr, 10) : 1;
      jgpm_csr_t h = nullptr;
      int st = jgpm_gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed, &h);
      if (st) { std::cerr << jgpm_strerror(st) << "\n"; return 2; }
      int32_t n = 0, md = 0;
      int64_t z = 0;
      jgpm_csr_info(h, &n, &z, &md);
      const int64_t* r = nullptr;
      const int32_t* c = nullptr;
      jgpm_csr_data(h, &r, &c);
      gp.reset(new Graph());
      gp->assign(std::vector<eidType>(r, r + n + 1), std::vector<vidType>(c, c + z));
      jgpm_csr_free(h);
      if (use_dag) gp->orientation();
    } else {
      gp.reset(new Graph(src, use_dag));
    }
    Graph& g = *gp;
    (void)dag_pattern;
    g.print_meta_data();
    uint64_t total = 0;
    int st = SubgraphSolver(g, total, argc, argv);
    return st == 0 ? 0 : 3;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 2;
  }
}
