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

VertexSet Graph::N(vidType v) const { return VertexSet(adj_ptr(v), get_degree(v), v); }

bool Graph::is_connected(std::vector<vidType> sg) const {
  if (sg.empty()) return false;
  std::sort(sg.begin(), sg.end());
  sg.erase(std::unique(sg.begin(), sg.end()), sg.end());
  std::vector<char> seen(sg.size(), 0);
  std::vector<size_t> stack{0};
  seen[0] = 1;
  size_t reached = 1;
  while (!stack.empty()) {
    size_t i = stack.back();
    stack.pop_back();
    for (size_t j = 0; j < sg.size(); ++j)
      if (!seen[j] && is_connected(sg[i], sg[j])) {
        seen[j] = 1;
        ++reached;
        stack.push_back(j);
      }
  }
  return reached == sg.size();
}

eidType Graph::init_edgelist(bool sym_break, bool) {
  if (nnz != 0) return nnz;
  nnz = sym_break ? n_edges / 2 : n_edges;
  src_list.assign((size_t)n_edges, 0);
  for (vidType v = 0; v < n_vertices; ++v)
    for (eidType e = vertices[v]; e < vertices[v + 1]; ++e) src_list[e] = v;
  return nnz;
}

void Graph::sort_graph() {
  // stable counting sort of the vertices by degree: order[k] is the vertex
  // that gets the new id k
  std::vector<eidType> cnt((size_t)max_degree + 2, 0);
  for (vidType v = 0; v < n_vertices; ++v) ++cnt[get_degree(v) + 1];
  for (size_t d = 1; d < cnt.size(); ++d) cnt[d] += cnt[d - 1];
  std::vector<vidType> order((size_t)n_vertices), rank((size_t)n_vertices);
  for (vidType v = 0; v < n_vertices; ++v) order[cnt[get_degree(v)]++] = v;
  for (vidType k = 0; k < n_vertices; ++k) rank[order[k]] = k;
  std::vector<eidType> rp((size_t)n_vertices + 1, 0);
  std::vector<vidType> ci((size_t)n_edges);
  for (vidType k = 0; k < n_vertices; ++k) rp[k + 1] = rp[k] + get_degree(order[k]);
  for (vidType k = 0; k < n_vertices; ++k) {
    const vidType v = order[k];
    eidType o = rp[k];
    for (eidType e = vertices[v]; e < vertices[v + 1]; ++e) ci[o++] = rank[edges[e]];
    std::sort(ci.begin() + rp[k], ci.begin() + rp[k + 1]);
  }
  vertices = std::move(rp);
  edges = std::move(ci);
  nnz = 0;
  src_list.clear();
}

void Graph::print_graph() const {
  std::cout << "Printing the graph: \n";
  for (vidType v = 0; v < n_vertices; ++v) {
    std::cout << "vertex " << v << ": degree = " << get_degree(v) << " edgelist = [ ";
    for (eidType e = edge_begin(v); e != edge_end(v); ++e) std::cout << getEdgeDst(e) << " ";
    std::cout << "]\n";
  }
}

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
  nnz = 0;
  src_list.clear();
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
