// Macro mesh and quadtree forest of the host setup (the mesh layer of
// cpp/include/insem/mesh/macro_mesh.hpp:42-106 and forest.hpp:50-121).
//
// Rules that fix the element order, and through it the assembly summation
// order of the hot path, and so are those of the reference: cells with a
// face on the axis are turned so that it is face 3; a tree's leaves are in
// Morton order and a refined leaf's four children replace it in place, in
// Morton order; 2:1 balance refines the coarser side; lattice vertices are
// numbered in order of first appearance over the leaves' corners.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

namespace insem_amd {

// Local faces as (first, second) corner, counter-clockwise: face 0 eta = -1,
// 1 xi = +1, 2 eta = +1, 3 xi = -1.
inline constexpr int kFaceCorner[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

struct BoundaryFace {
  int32_t cell;
  int face;
  int tag;
};

class MacroMesh {
 public:
  // nodes_per_cell is 4 (bilinear) or 9 (biquadratic); axisymmetric meshes
  // use (r, z) with the axis at r = 0.
  MacroMesh(int nodes_per_cell, bool axisymmetric, std::vector<double> nodes, std::vector<int32_t> cells,
            std::vector<int> material, std::vector<BoundaryFace> boundary);

  int nodes_per_cell() const { return npc_; }
  bool axisymmetric() const { return axisym_; }
  int32_t n_nodes() const { return static_cast<int32_t>(nodes_.size() / 2); }
  int32_t n_cells() const { return static_cast<int32_t>(material_.size()); }
  int32_t n_edges() const { return static_cast<int32_t>(edge_cells_.size() / 2); }
  int32_t cell_node(int32_t c, int k) const { return cells_[static_cast<std::size_t>(c) * npc_ + k]; }
  int material(int32_t c) const { return material_[c]; }
  const std::vector<int>& materials() const { return material_; }
  const std::vector<double>& nodes() const { return nodes_; }
  const std::vector<int32_t>& cells() const { return cells_; }
  const std::vector<BoundaryFace>& boundary() const { return boundary_; }

  int32_t cell_edge(int32_t c, int f) const { return cell_edges_[4 * static_cast<std::size_t>(c) + f]; }
  bool cell_edge_forward(int32_t c, int f) const { return edge_fwd_[4 * static_cast<std::size_t>(c) + f] != 0; }
  std::array<int32_t, 2> edge_cells(int32_t e) const { return {edge_cells_[2 * e], edge_cells_[2 * e + 1]}; }
  bool is_boundary_edge(int32_t e) const { return edge_cells_[2 * e + 1] < 0; }

  bool is_axis_face(int32_t c, int f) const;
  bool touches_axis(int32_t c) const { return is_axis_face(c, 3); }

  std::array<double, 2> map(int32_t c, double xi, double eta) const;
  // {dx/dxi, dx/deta, dz/dxi, dz/deta}
  std::array<double, 4> jacobian(int32_t c, double xi, double eta) const;

 private:
  void rotate_axis_cells();
  void build_edges();
  void check() const;

  int npc_;
  bool axisym_;
  std::vector<double> nodes_;
  std::vector<int32_t> cells_;
  std::vector<int> material_;
  std::vector<BoundaryFace> boundary_;
  std::vector<int32_t> cell_edges_;
  std::vector<uint8_t> edge_fwd_;
  std::vector<int32_t> edge_cells_;
};

struct Leaf {
  int32_t tree;
  int32_t level;
  int32_t i;
  int32_t j;
};

struct Neighbor {
  int32_t leaf = -1;  // -1 on the domain boundary
  int face = -1;      // the neighbour's face
  int level_diff = 0; // neighbour level - own level
};

struct Corners {
  int32_t n_vertices = 0;
  std::vector<int32_t> corner;  // [4 * leaf + k]
};

class Forest {
 public:
  static constexpr int kMaxLevel = 20;

  static Forest uniform(std::shared_ptr<const MacroMesh> macro, int level);

  const MacroMesh& macro() const { return *macro_; }
  std::shared_ptr<const MacroMesh> macro_ptr() const { return macro_; }
  int32_t n_leaves() const { return static_cast<int32_t>(leaves_.size()); }
  const Leaf& leaf(int32_t l) const { return leaves_[l]; }
  std::array<double, 4> leaf_box(int32_t l) const;
  Corners corners() const;
  Neighbor neighbor(int32_t l, int f) const;
  double face_coordinate(int32_t from, int f, double t, int32_t to, int to_face) const;
  Forest refine(const std::vector<char>& mark) const;
  Forest balance(const std::function<bool(int32_t, int32_t)>& strict_pair = {}) const;
  int max_level() const;
  bool is_axial(int32_t l) const { return leaves_[l].i == 0 && macro_->touches_axis(leaves_[l].tree); }

 private:
  Forest(std::shared_ptr<const MacroMesh> macro, std::vector<Leaf> leaves);
  // the leaf of a tree that covers max-level cell (x, y), or -1
  int32_t leaf_at(int32_t tree, int64_t x, int64_t y) const;
  std::shared_ptr<const MacroMesh> macro_;
  std::vector<Leaf> leaves_;
  // leaves are sorted by (tree, Morton code of their max-level anchor cell)
  std::vector<uint64_t> code_;        // Morton code of each leaf's anchor
  std::vector<int32_t> tree_first_;   // first leaf of each tree (+ end)
};

}  // namespace insem_amd
