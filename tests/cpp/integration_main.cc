// main.cc (reference layout), one-GPU build: the caller-side program of
// INTEGRATION.md, built against include/ and libjgpm.so by
// tests/test_host_api.py.
#include "graph.h"            // this repository's include/graph.h
int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cout << "Usage: " << argv[0] << " <graph> [pattern(P1)] [round(1)]\n";
    return 1;
  }
  std::string p(argv[2]);
  bool use_dag = p == "P13" || p == "P14" || p == "P15" || p == "P16";
  Graph g(argv[1], use_dag);                     // same loader, same file format
  g.print_meta_data();
  uint64_t total = 0;
  return SubgraphSolver(g, total, argc, argv);   // fills `total` (the reference never did)
}
