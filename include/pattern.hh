// Pattern — small query graph read from an edge-list file (one "u v" pair
// per line), with the naming of the reference's Pattern (include/pattern.hh,
// common/pattern.cc:3-27,80-137).  pattern_id() maps a pattern graph to the
// library's pattern id (TC, P1..P16 and the 3-star / tailed-triangle
// extensions) by isomorphism, or -1 if it is not in the catalogue.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "VertexSet.h"

class Pattern {
 public:
  Pattern() {}
  explicit Pattern(std::string name) : name_(name) {}
  Pattern(std::string filename, bool is_labeled) { read_adj_file(filename); set_name(); }

  void read_adj_file(std::string inputfile);
  void set_name();
  std::string get_name() const { return name_; }
  int size() const { return num_vertices; }
  int sizeEdges() const { return num_edges; }
  int get_num_vertices() const { return num_vertices; }
  int get_num_edges() const { return num_edges; }
  int get_degree(vidType v) const { return (int)adj_list.at(v).size(); }
  int get_max_degree() const { return max_degree; }
  const std::vector<vidType>& get_neighbours(vidType v) const { return adj_list.at(v); }
  const std::vector<vidType>& N(vidType v) const { return adj_list.at(v); }
  // vertex -> neighbour list of the pattern (for plan compilation)
  const std::map<vidType, std::vector<vidType>>& adjacency() const { return adj_list; }
  bool is_connected(vidType u, vidType v) const;
  Pattern& add_edge(vidType u, vidType v);
  Pattern& remove_edge(vidType u, vidType v);
  std::string to_string() const;
  // name tests of the reference (include/pattern.hh:55-61); set_name gives
  // the reference's names (common/pattern.cc:3-27), so is_rectangle,
  // is_pentagon, is_house and is_4color_square stay false there as here
  bool is_wedge() const { return name_ == "wedge"; }
  bool is_triangle() const { return name_ == "triangle"; }
  bool is_diamond() const { return name_ == "diamond"; }
  bool is_rectangle() const { return name_ == "rectangle"; }
  bool is_pentagon() const { return name_ == "pentagon"; }
  bool is_house() const { return name_ == "house"; }
  bool is_4color_square() const { return name_ == "4color-square"; }
  vidType get_neighbor(vidType v, vidType i) const { return adj_list.at(v)[i]; }
  // pattern vertex ids, ascending
  std::vector<vidType> v_list() const;
  void print_meta_data() const;
  int pattern_id() const;

 private:
  void recount();
  std::string name_;
  std::map<vidType, std::vector<vidType>> adj_list;
  int num_vertices = 0, num_edges = 0, max_degree = 0;
};
