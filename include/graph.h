// Graph — host CSR graph with the public interface of the reference's Graph
// (include/graph.h:13-149): load <prefix>.meta.txt/.vertex.bin/.edge.bin,
// optional degree-DAG orientation, CSR accessors and VertexSet views.  The
// arrays are handed to libjgpm (include/jgpm.h) for device counting.
#pragma once
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
  vidType* adj_ptr(vidType v) { return &edges[vertices[v]]; }
  vidType N(vidType v, vidType n) const { return edges[vertices[v] + n]; }
  VertexSet N(vidType v);
  eidType* out_rowptr() { return vertices.data(); }
  vidType* out_colidx() { return edges.data(); }
  bool is_connected(vidType v, vidType u) const;
  vidType getEdgeDst(eidType e) const { return edges[e]; }
  vidType size() const { return n_vertices; }
  eidType sizeEdges() const { return n_edges; }

  void fixEndEdge(vidType vid, eidType row_end) { vertices[vid + 1] = row_end; }
  void allocateFrom(vidType nv, eidType ne);
  void constructEdge(eidType eid, vidType dst) { edges[eid] = dst; }

  // Degree-ordered DAG: keep u->v iff d(v) > d(u) or (d(v) == d(u) and v > u).
  void orientation();
  void dump_binary(std::string prefix);
  void print_meta_data() const;
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
};

// One-GPU drop-in for the reference driver (src/jupiter_gpm.cu:27): argv[2]
// is the pattern name, argv[3] (optional) the number of rounds the root
// range is split into.  Unlike the reference, `total` receives the count.
// Returns 0 or a negative jgpm status.
int SubgraphSolver(Graph& g, uint64_t& total, int argc, char* argv[]);
