// Exercises the reference-shaped host API members of include/graph.h and
// include/pattern.hh (no device calls).  Prints one "key value..." line per
// check; tests/test_host_api.py compares them with numpy.
#include <cstdio>
#include <iostream>
#include <vector>

#include "graph.h"
#include "pattern.hh"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  // 6 vertices: a triangle 0-1-2, a path 2-3-4, and an isolated vertex 5
  Graph g;
  g.assign({0, 2, 4, 7, 9, 10, 10}, {1, 2, 0, 2, 0, 1, 3, 2, 4, 3});
  std::printf("tasks_before %lld\n", (long long)g.get_num_tasks());
  eidType nt = g.init_edgelist();
  std::printf("tasks %lld %lld\n", (long long)nt, (long long)g.get_num_tasks());
  std::printf("src");
  for (eidType e = 0; e < g.E(); ++e) std::printf(" %d", g.get_src(e));
  std::printf("\ndst");
  for (eidType e = 0; e < g.E(); ++e) std::printf(" %d", g.get_dst(e));
  std::printf("\nsrc_ptr_ok %d\n", g.get_src_ptr()[3] == g.get_src(3) && g.get_dst_ptr()[3] == g.get_dst(3));
  std::printf("connected %d %d %d %d\n", g.is_connected(std::vector<vidType>{0, 1, 2, 3, 4}),
              g.is_connected(std::vector<vidType>{0, 3}), g.is_connected(std::vector<vidType>{4, 5}),
              g.is_connected(std::vector<vidType>{3}));
  const Graph& cg = g;
  VertexSet n2 = cg.N(2);
  std::printf("N2 %d:", n2.size());
  for (vidType i = 0; i < n2.size(); ++i) std::printf(" %d", n2[i]);
  std::printf("\n");
  g.sort_graph();
  std::printf("sorted_rowptr");
  for (vidType v = 0; v < g.V(); ++v) std::printf(" %lld", (long long)g.edge_begin(v));
  std::printf(" %lld", (long long)g.edge_end(g.V() - 1));
  std::printf("\nsorted_col");
  for (eidType e = 0; e < g.E(); ++e) std::printf(" %d", g.getEdgeDst(e));
  std::printf("\n");
  std::fflush(stdout);
  g.print_graph();
  std::cout.flush();
  Pattern p(argv[1], false);
  std::printf("pattern %s wedge %d triangle %d diamond %d rect %d pent %d house %d sq4 %d\n", p.get_name().c_str(),
              p.is_wedge(), p.is_triangle(), p.is_diamond(), p.is_rectangle(), p.is_pentagon(), p.is_house(),
              p.is_4color_square());
  std::printf("v_list");
  for (vidType v : p.v_list()) std::printf(" %d", v);
  std::printf("\nneighbor %d %d\n", p.get_neighbor(p.v_list()[0], 0), p.pattern_id());
  std::fflush(stdout);
  p.print_meta_data();
  Pattern q(argv[2], false);
  std::printf("pattern2 %s %d\n", q.get_name().c_str(), q.pattern_id());
  return 0;
}
