// Graph — host CSR graph with the public interface of the reference's Graph
// (include/graph.h:13-149): load <prefix>.meta.txt/.vertex.bin/.edge.bin,
// optional degree-DAG orientation, CSR accessors and VertexSet views.  The
// arrays are handed to libjgpm (include/jgpm.h) for device counting.
#pragma once
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

#include "VertexSet.h"

class Graph {
 public:
  Graph(std::string prefix, bool use_dag = false, bool directed = false, bool use_vlabel = false,
        bool use_elabel = false);
  Graph() {}
  Graph(vidType nv, eidType ne) { allocateFrom(nv, ne); }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  vidType V() const { return n_vertices; }
  eidType E() const { return n_edges; }
  // number of COO tasks built by init_edgelist (0 before)
  eidType get_num_tasks() const { return nnz; }
  vidType num_vertices() const { return n_vertices; }
  eidType num_edges() const { return n_edges; }
  std::string get_name() const { return name; }
  std::string get_inputfile_path() const { return inputfile_path; }
  bool is_directed() const { return is_directed_; }
  bool is_dag() const { return is_dag_; }
  void set_max_degree(vidType md) { max_degree = md; }
  vidType get_max_degree() const { return max_degree; }

  vidType get_degree(vidType v) const { return (vidType)(vertices[v + 1] - vertices[v]); }
  vidType out_degree(vidType v) const { return get_degree(v); }
  eidType edge_begin(vidType v) const { return vertices[v]; }
  eidType edge_end(vidType v) const { return vertices[v + 1]; }
  vidType* adj_ptr(vidType v) const { return const_cast<vidType*>(&edges[vertices[v]]); }
  vidType N(vidType v, vidType n) const { return edges[vertices[v] + n]; }
  VertexSet N(vidType v) const;
  eidType* out_rowptr() { return vertices.data(); }
  vidType* out_colidx() { return edges.data(); }
  bool is_connected(vidType v, vidType u) const;
  // true iff the subgraph induced by the vertices of sg is connected (the
  // reference declares this member but its body returns false,
  // common/graph.cc:525-528)
  bool is_connected(std::vector<vidType> sg) const;
  vidType getEdgeDst(eidType e) const { return edges[e]; }
  vidType size() const { return n_vertices; }
  eidType sizeEdges() const { return n_edges; }

  void fixEndEdge(vidType vid, eidType row_end) { vertices[vid + 1] = row_end; }
  void allocateFrom(vidType nv, eidType ne);
  void constructEdge(eidType eid, vidType dst) { edges[eid] = dst; }

  // COO edge list (common/graph.cc:462-506): src[e] is the row of edge e,
  // dst[e] = colidx[e].  Returns the number of tasks (E(), halved when
  // sym_break, as the reference sizes it).
  eidType init_edgelist(bool sym_break = false, bool ascend = false);
  vidType* get_src_ptr() { return src_list.data(); }
  vidType* get_dst_ptr() { return edges.data(); }
  vidType get_src(eidType eid) { return src_list[eid]; }
  vidType get_dst(eidType eid) { return edges[eid]; }
  std::vector<vidType> get_sizes() const { return sizes; }

  // Degree-ordered DAG: keep u->v iff d(v) > d(u) or (d(v) == d(u) and v > u).
  void orientation();
  // Relabel the vertices by ascending degree (ties by ascending id, the
  // stable counting sort of common/graph.cc:602-725) and sort every list.
  void sort_graph();
  void dump_binary(std::string prefix);
  void print_meta_data() const;
  void print_graph() const;
  // Adopt a CSR (e.g. from a generator); lists must be ascending.
  void assign(std::vector<eidType> rowptr, std::vector<vidType> colidx);

 private:
  std::string name, inputfile_path;
  bool is_directed_ = false;
  bool is_dag_ = false;
  vidType max_degree = 0;
  vidType n_vertices = 0;
  eidType n_edges = 0;
  std::vector<eidType> vertices;
  std::vector<vidType> edges;
  eidType nnz = 0;                 // COO tasks (init_edgelist)
  std::vector<vidType> src_list;   // COO sources
  std::vector<vidType> sizes;      // per-vertex task counts (kept empty, as the reference leaves it)
};

// One-GPU drop-in for the reference driver (src/jupiter_gpm.cu:27): argv[2]
// is the pattern name, argv[3] (optional) the number of rounds the root
// range is split into.  Unlike the reference, `total` receives the count.
// Returns 0 or a negative jgpm status.
int SubgraphSolver(Graph& g, uint64_t& total, int argc, char* argv[]);
// Pattern id of a pattern name ("P13", "TC", "I_C4", ...), or -1.
int jgpm_pattern_from_name(const std::string& name);
