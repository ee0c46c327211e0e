// jupiter_gpm — drop-in command line of the reference (src/main.cc:7-48):
//   jupiter_gpm <graph_prefix> <pattern> [total_round=1]
// <pattern> as SubgraphSolver takes it (include/graph.h).  <graph_prefix>
// names the .meta.txt/.vertex.bin/.edge.bin files; "rmat:<scale>[:<seed>]"
// generates a seeded R-MAT graph (edge factor 16, a/b/c = 0.57/0.19/0.19)
// instead.  P13..P16 load the degree DAG, as in the reference.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "graph.h"
#include "jgpm.h"

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cout << "Usage: " << argv[0] << " <graph> [pattern(P1)] [round(1)]\n";
    std::cout << "Example: " << argv[0] << " /graph_inputs/mico/graph P1\n";
    return 1;
  }
  std::string pname(argv[2]);
  int pid = jgpm_pattern_from_name(pname);
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
    g.print_meta_data();
    uint64_t total = 0;
    int st = SubgraphSolver(g, total, argc, argv);
    return st == 0 ? 0 : 3;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 2;
  }
}
