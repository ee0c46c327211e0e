// Pattern graphs (include/pattern.hh): edge-list reader, naming, and the
// isomorphism lookup into the library's pattern catalogue.
#include <algorithm>
#include <fstream>
#include <functional>
#include <iostream>
#include <iterator>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "jgpm.h"
#include "pattern.hh"

void Pattern::read_adj_file(std::string inputfile) {
  std::ifstream in(inputfile.c_str());
  if (!in) throw std::invalid_argument("cannot open pattern file '" + inputfile + "'");
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::vector<vidType> vs{std::istream_iterator<vidType>(iss), std::istream_iterator<vidType>()};
    if (vs.empty()) continue;
    if (vs.size() == 2) add_edge(vs[0], vs[1]);
    else if (vs.size() == 4) add_edge(vs[0], vs[2]);  // labelled edge "u lu v lv": labels ignored
    else throw std::invalid_argument("Input file '" + inputfile + "' error format");
  }
  if (adj_list.empty()) throw std::invalid_argument("Found 0 vertices in file '" + inputfile + "'");
}

void Pattern::recount() {
  num_vertices = (int)adj_list.size();
  num_edges = 0;
  max_degree = 0;
  for (auto& kv : adj_list) {
    num_edges += (int)kv.second.size();
    max_degree = std::max<int>(max_degree, (int)kv.second.size());
  }
  num_edges /= 2;
}

Pattern& Pattern::add_edge(vidType u, vidType v) {
  if (u == v || is_connected(u, v)) return *this;
  adj_list[u].push_back(v);
  adj_list[v].push_back(u);
  recount();
  return *this;
}

Pattern& Pattern::remove_edge(vidType u, vidType v) {
  auto rm = [&](vidType a, vidType b) {
    auto it = adj_list.find(a);
    if (it == adj_list.end()) return;
    auto& l = it->second;
    l.erase(std::remove(l.begin(), l.end(), b), l.end());
  };
  rm(u, v);
  rm(v, u);
  recount();
  return *this;
}

bool Pattern::is_connected(vidType u, vidType v) const {
  auto it = adj_list.find(u);
  if (it == adj_list.end()) return false;
  return std::find(it->second.begin(), it->second.end(), v) != it->second.end();
}

// Names by vertex and edge count only, as the reference does: every
// 3-edge 4-vertex pattern is "3-star" and every 4-edge one "square"
// (common/pattern.cc:3-27); pattern_id() tells the shapes apart.
void Pattern::set_name() {
  const int n = num_vertices, m = num_edges;
  name_ = "unknown";
  if (n == 3) name_ = m == 2 ? "wedge" : "triangle";
  if (n == 4) {
    if (m == 3) name_ = "3-star";
    if (m == 4) name_ = "square";
    if (m == 5) name_ = "diamond";
    if (m == 6) name_ = "4-clique";
  }
}

std::vector<vidType> Pattern::v_list() const {
  std::vector<vidType> vs;
  for (auto& kv : adj_list) vs.push_back(kv.first);  // std::map: ascending
  return vs;
}

void Pattern::print_meta_data() const {
  std::cout << "|V|: " << num_vertices << ", |E|: " << num_edges << ", |\u03A3|: 0" << std::endl;
  std::cout << "Max Degree: " << max_degree << ", Max Label Frequency: 0" << std::endl;
}

std::string Pattern::to_string() const {
  std::string s;
  for (auto& kv : adj_list)
    for (vidType v : kv.second)
      if (kv.first < v) s += "[" + std::to_string(kv.first) + "-" + std::to_string(v) + "]";
  return s;
}

namespace {
// Edge lists of the catalogue (vertex 0..k-1): each P-id's pattern graph H.
struct Shape { int id; int k; std::vector<std::pair<int, int>> e; };
std::vector<Shape> catalogue() {
  auto clique = [](int k) {
    std::vector<std::pair<int, int>> e;
    for (int a = 0; a < k; ++a) for (int b = a + 1; b < k; ++b) e.push_back({a, b});
    return e;
  };
  return {
      {JGPM_TC, 3, clique(3)},
      {JGPM_P1, 4, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}},
      {JGPM_P2, 4, {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}}},
      {JGPM_P3, 5, {{0, 1}, {0, 2}, {1, 2}, {0, 3}, {3, 4}}},
      {JGPM_P4, 5, {{0, 1}, {0, 2}, {1, 2}, {1, 3}, {3, 4}, {4, 0}}},
      {JGPM_P5, 5, {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}, {4, 0}, {4, 1}}},
      {JGPM_P6, 5, {{0, 1}, {2, 0}, {2, 1}, {3, 0}, {3, 1}, {4, 2}, {4, 3}}},
      {JGPM_P7, 5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {4, 1}, {1, 2}, {2, 3}}},
      {JGPM_P8, 5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 2}, {2, 3}, {3, 4}, {4, 1}}},
      {JGPM_P9, 5, {{0, 1}, {0, 2}, {1, 2}, {3, 0}, {3, 1}, {3, 2}, {4, 0}, {4, 1}, {4, 2}}},
      {JGPM_P10, 6, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {4, 1}, {1, 2}, {2, 3}, {3, 5}}},
      {JGPM_P11, 6, {{0, 1}, {0, 2}, {1, 2}, {3, 0}, {3, 1}, {4, 0}, {4, 2}, {5, 1}, {5, 2}}},
      {JGPM_P12, 6, {{0, 1}, {2, 0}, {2, 1}, {3, 0}, {3, 1}, {4, 0}, {4, 1}, {5, 0}, {5, 1}, {2, 3}, {2, 4}}},
      {JGPM_P13, 4, clique(4)}, {JGPM_P14, 5, clique(5)}, {JGPM_P15, 6, clique(6)}, {JGPM_P16, 7, clique(7)},
      {JGPM_STAR3, 4, {{0, 1}, {0, 2}, {0, 3}}},
      {JGPM_TAILTRI, 4, {{0, 1}, {0, 2}, {1, 2}, {0, 3}}},
  };
}
}  // namespace

int Pattern::pattern_id() const {
  // relabel to 0..n-1 and test every permutation against each shape
  std::vector<vidType> ids;
  for (auto& kv : adj_list) ids.push_back(kv.first);
  const int n = (int)ids.size();
  if (n > 7) return -1;
  std::vector<std::vector<char>> A(n, std::vector<char>(n, 0));
  for (int a = 0; a < n; ++a)
    for (vidType v : adj_list.at(ids[a])) {
      int b = (int)(std::find(ids.begin(), ids.end(), v) - ids.begin());
      A[a][b] = A[b][a] = 1;
    }
  for (const Shape& s : catalogue()) {
    if (s.k != n || (int)s.e.size() != num_edges) continue;
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    do {
      bool ok = true;
      for (auto& e : s.e) ok &= A[perm[e.first]][perm[e.second]] == 1;
      if (ok) return s.id;
    } while (std::next_permutation(perm.begin(), perm.end()));
  }
  return -1;
}
