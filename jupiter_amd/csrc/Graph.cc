// Host Graph (include/graph.h): file loading and orientation go through the
// host CSR functions of libjgpm (include/jgpm.h); the counting entry point
// SubgraphSolver drives the device through the same C ABI.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "graph.h"
#include "jgpm.h"

static void take(jgpm_csr_t h, std::vector<eidType>& rp, std::vector<vidType>& ci, vidType& nv, eidType& ne,
                 vidType& md) {
  int32_t n = 0, m = 0;
  int64_t z = 0;
  jgpm_csr_info(h, &n, &z, &m);
  const int64_t* r = nullptr;
  const int32_t* c = nullptr;
  jgpm_csr_data(h, &r, &c);
  rp.assign(r, r + n + 1);
  ci.assign(c, c + z);
  nv = n;
  ne = z;
  md = m;
  jgpm_csr_free(h);
}

Graph::Graph(std::string prefix, bool use_dag, bool directed, bool, bool) : is_directed_(directed) {
  size_t i = prefix.rfind('/');
  if (i != std::string::npos) inputfile_path = prefix.substr(0, i);
  size_t j = inputfile_path.rfind('/');
  name = j != std::string::npos ? inputfile_path.substr(j + 1) : inputfile_path;
  jgpm_csr_t h = nullptr;
  int st = jgpm_csr_read(prefix.c_str(), &h);
  if (st) throw std::runtime_error("cannot load graph " + prefix + ": " + jgpm_strerror(st));
  take(h, vertices, edges, n_vertices, n_edges, max_degree);
  if (use_dag) orientation();
  if (max_degree > VertexSet::MAX_DEGREE) VertexSet::MAX_DEGREE = max_degree;
}

void Graph::allocateFrom(vidType nv, eidType ne) {
  n_vertices = nv;
  n_edges = ne;
  vertices.assign((size_t)nv + 1, 0);
  edges.assign((size_t)ne, 0);
}

void Graph::assign(std::vector<eidType> rowptr, std::vector<vidType> colidx) {
  vertices = std::move(rowptr);
  edges = std::move(colidx);
  n_vertices = (vidType)vertices.size() - 1;
  n_edges = (eidType)edges.size();
  max_degree = 0;
  for (vidType v = 0; v < n_vertices; ++v) max_degree = std::max(max_degree, get_degree(v));
  if (max_degree > VertexSet::MAX_DEGREE) VertexSet::MAX_DEGREE = max_degree;
}

VertexSet Graph::N(vidType v) { return VertexSet(edges.data() + vertices[v], get_degree(v), v); }

bool Graph::is_connected(vidType v, vidType u) const {
  if (get_degree(v) > get_degree(u)) std::swap(u, v);
  const vidType* b = edges.data() + vertices[v];
  const vidType* e = edges.data() + vertices[v + 1];
  return std::binary_search(b, e, u);
}

void Graph::orientation() {
  // the symmetric CSR goes through the library's orientation routine
  std::vector<int32_t> src((size_t)n_edges);
  for (vidType v = 0; v < n_vertices; ++v)
    for (eidType e = vertices[v]; e < vertices[v + 1]; ++e) src[e] = v;
  jgpm_csr_t sym = nullptr, dag = nullptr;
  int st = jgpm_csr_from_edges(n_vertices, n_edges, src.data(), edges.data(), &sym);
  if (st) throw std::runtime_error(std::string("orientation: ") + jgpm_strerror(st));
  st = jgpm_csr_orient(sym, &dag);
  jgpm_csr_free(sym);
  if (st) throw std::runtime_error(std::string("orientation: ") + jgpm_strerror(st));
  take(dag, vertices, edges, n_vertices, n_edges, max_degree);
  is_dag_ = true;
}

void Graph::dump_binary(std::string prefix) {
  std::vector<int32_t> src((size_t)n_edges);
  for (vidType v = 0; v < n_vertices; ++v)
    for (eidType e = vertices[v]; e < vertices[v + 1]; ++e) src[e] = v;
  // write the arrays verbatim (a DAG stays a DAG)
  FILE* fv = fopen((prefix + ".vertex.bin").c_str(), "wb");
  FILE* fe = fopen((prefix + ".edge.bin").c_str(), "wb");
  FILE* fm = fopen((prefix + ".meta.txt").c_str(), "w");
  if (!fv || !fe || !fm) throw std::runtime_error("dump_binary: cannot open " + prefix);
  fwrite(vertices.data(), sizeof(eidType), vertices.size(), fv);
  fwrite(edges.data(), sizeof(vidType), edges.size(), fe);
  fprintf(fm, "%d\n%ld\n4 8 1 2\n%d\n0\n0\n0\n", n_vertices, (long)n_edges, max_degree);
  fclose(fv);
  fclose(fe);
  fclose(fm);
}

void Graph::print_meta_data() const {
  std::cout << "|V|: " << n_vertices << ", |E|: " << n_edges << ", Max Degree: " << max_degree << "\n";
}
