#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <utility>

#include "error.hpp"

namespace insem_amd {
namespace {

// Node k of a cell on the 3 x 3 lattice of its reference square (0, 1, 2 =
// -1, 0, +1 in xi and eta): corners counter-clockwise, then the midpoints of
// faces 0..3, then the centre.
constexpr int kNodeXi[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
constexpr int kNodeEta[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};

// 1D Lagrange bases of the cell shape functions on the nodes of kNodeXi:
// linear on {-1, +1} (lattice 0 and 2), quadratic on {-1, 0, +1}.
void linear_1d(double t, double* l, double* dl) {
  l[0] = 0.5 * (1.0 - t), l[2] = 0.5 * (1.0 + t);
  dl[0] = -0.5, dl[2] = 0.5;
}

void quadratic_1d(double t, double* l, double* dl) {
  l[0] = 0.5 * t * (t - 1.0), l[1] = 1.0 - t * t, l[2] = 0.5 * t * (t + 1.0);
  dl[0] = t - 0.5, dl[1] = -2.0 * t, dl[2] = t + 0.5;
}

// Shape functions of a 4- or 9-node cell and their xi, eta derivatives at
// (xi, eta), in node order.
struct CellShape {
  int n = 0;
  double value[9], d_xi[9], d_eta[9];

  CellShape(int npc, double xi, double eta) : n(npc) {
    double lx[3] = {}, dlx[3] = {}, ly[3] = {}, dly[3] = {};
    if (npc == 4) {
      linear_1d(xi, lx, dlx);
      linear_1d(eta, ly, dly);
    } else {
      quadratic_1d(xi, lx, dlx);
      quadratic_1d(eta, ly, dly);
    }
    for (int k = 0; k < n; ++k) {
      value[k] = lx[kNodeXi[k]] * ly[kNodeEta[k]];
      d_xi[k] = dlx[kNodeXi[k]] * ly[kNodeEta[k]];
      d_eta[k] = lx[kNodeXi[k]] * dly[kNodeEta[k]];
    }
  }
};

}  // namespace

MacroMesh::MacroMesh(int nodes_per_cell, bool axisymmetric, std::vector<double> nodes, std::vector<int32_t> cells,
                     std::vector<int> material, std::vector<BoundaryFace> boundary)
    : npc_(nodes_per_cell),
      axisym_(axisymmetric),
      nodes_(std::move(nodes)),
      cells_(std::move(cells)),
      material_(std::move(material)),
      boundary_(std::move(boundary)) {
  INSEM_REQUIRE(npc_ == 4 || npc_ == 9, "a macro cell has 4 or 9 nodes, not " << npc_);
  INSEM_REQUIRE(nodes_.size() % 2 == 0, "node coordinates come in (x, z) pairs");
  INSEM_REQUIRE(cells_.size() == material_.size() * static_cast<std::size_t>(npc_),
                material_.size() << " cells of " << npc_ << " nodes need " << material_.size() * npc_
                                 << " node indices, not " << cells_.size());
  for (int32_t v : cells_) INSEM_REQUIRE(v >= 0 && v < n_nodes(), "a cell names node " << v << ", which does not exist");
  if (axisym_) rotate_axis_cells();
  build_edges();
  check();
}

// A face lies on the axis when both its corners have r within 1e-10 of the
// cell's extent.
bool MacroMesh::is_axis_face(int32_t c, int f) const {
  if (!axisym_) return false;
  double extent = 0.0;
  for (int k = 0; k < 4; ++k) {
    const double* p = &nodes_[2 * static_cast<std::size_t>(cell_node(c, k))];
    extent = std::max(extent, std::max(std::fabs(p[0]), std::fabs(p[1])));
  }
  const double tol = 1e-10 * std::max(extent, 1e-300);
  for (int s = 0; s < 2; ++s)
    if (std::fabs(nodes_[2 * static_cast<std::size_t>(cell_node(c, kFaceCorner[f][s]))]) > tol) return false;
  return true;
}

// An axisymmetric cell with a face on the axis is turned (its corner and
// face-midpoint lists shifted) so that this face becomes face 3, the one
// the GLJ rule runs along.
void MacroMesh::rotate_axis_cells() {
  for (int32_t c = 0; c < n_cells(); ++c) {
    int on_axis = -1;
    for (int f = 0; f < 4; ++f)
      if (is_axis_face(c, f)) {
        INSEM_REQUIRE(on_axis < 0, "cell " << c << " touches the axis with two faces");
        on_axis = f;
      }
    if (on_axis < 0 || on_axis == 3) continue;
    const int shift = (on_axis + 1) % 4;  // new corner i = old corner i + shift
    int32_t* node = &cells_[static_cast<std::size_t>(c) * npc_];
    std::rotate(node, node + shift, node + 4);
    if (npc_ == 9) std::rotate(node + 4, node + 4 + shift, node + 8);
    for (BoundaryFace& bf : boundary_)
      if (bf.cell == c) bf.face = (bf.face + 4 - shift) % 4;
  }
}

// Macro edges from the cells' faces (in cell and face order); an edge runs
// from its lower node index to its higher, and a cell face is forward when
// it runs the same way.
void MacroMesh::build_edges() {
  const std::size_t nc = static_cast<std::size_t>(n_cells());
  cell_edges_.assign(4 * nc, -1);
  edge_fwd_.assign(4 * nc, 0);
  edge_cells_.clear();
  // faces sorted by their node pair meet their partner next to them
  struct FaceRef {
    int32_t lo, hi, cell;
    int face;
  };
  std::vector<FaceRef> faces;
  faces.reserve(4 * nc);
  for (int32_t c = 0; c < n_cells(); ++c)
    for (int f = 0; f < 4; ++f) {
      const int32_t a = cell_node(c, kFaceCorner[f][0]), b = cell_node(c, kFaceCorner[f][1]);
      INSEM_REQUIRE(a != b, "face " << f << " of cell " << c << " joins node " << a << " to itself");
      faces.push_back({std::min(a, b), std::max(a, b), c, f});
      edge_fwd_[4 * static_cast<std::size_t>(c) + f] = a < b ? 1 : 0;
    }
  std::vector<std::size_t> order(faces.size());
  for (std::size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) {
    return faces[x].lo != faces[y].lo ? faces[x].lo < faces[y].lo : faces[x].hi < faces[y].hi;
  });
  for (std::size_t k = 0; k < order.size();) {
    std::size_t end = k + 1;
    while (end < order.size() && faces[order[end]].lo == faces[order[k]].lo && faces[order[end]].hi == faces[order[k]].hi)
      ++end;
    const FaceRef& a = faces[order[k]];
    INSEM_REQUIRE(end - k <= 2, "macro edge (" << a.lo << ", " << a.hi << ") belongs to " << end - k << " cells");
    const int32_t e = n_edges();
    edge_cells_.push_back(a.cell);
    edge_cells_.push_back(end - k == 2 ? faces[order[k + 1]].cell : -1);
    cell_edges_[4 * static_cast<std::size_t>(a.cell) + a.face] = e;
    if (end - k == 2) {
      const FaceRef& b = faces[order[k + 1]];
      cell_edges_[4 * static_cast<std::size_t>(b.cell) + b.face] = e;
      // two counter-clockwise cells traverse a shared edge in opposite directions
      INSEM_REQUIRE(cell_edge_forward(a.cell, a.face) != cell_edge_forward(b.cell, b.face),
                    "cells " << a.cell << " and " << b.cell << " are not both counter-clockwise");
    }
    k = end;
  }
}

void MacroMesh::check() const {
  for (int32_t c = 0; c < n_cells(); ++c)
    for (int q = 0; q < 9; ++q) {
      const auto j = jacobian(c, kNodeXi[q] - 1.0, kNodeEta[q] - 1.0);
      const double det = j[0] * j[3] - j[1] * j[2];
      INSEM_REQUIRE(det > 0.0, "cell " << c << " is inverted or degenerate (Jacobian " << det
                                       << "); list its corners counter-clockwise");
    }
  if (axisym_) {
    double extent = 0.0;
    for (double v : nodes_) extent = std::max(extent, std::fabs(v));
    for (int32_t i = 0; i < n_nodes(); ++i)
      INSEM_REQUIRE(nodes_[2 * static_cast<std::size_t>(i)] >= -1e-12 * extent,
                    "node " << i << " lies at r < 0 in an axisymmetric mesh");
  }
  for (const BoundaryFace& bf : boundary_) {
    INSEM_REQUIRE(bf.cell >= 0 && bf.cell < n_cells() && bf.face >= 0 && bf.face < 4,
                  "boundary entry (" << bf.cell << ", " << bf.face << ") names no cell face");
    INSEM_REQUIRE(is_boundary_edge(cell_edge(bf.cell, bf.face)),
                  "face " << bf.face << " of cell " << bf.cell << " is interior but carries a boundary tag");
  }
}

std::array<double, 2> MacroMesh::map(int32_t c, double xi, double eta) const {
  const CellShape s(npc_, xi, eta);
  std::array<double, 2> x{0.0, 0.0};
  for (int k = 0; k < s.n; ++k) {
    const double* p = &nodes_[2 * static_cast<std::size_t>(cell_node(c, k))];
    x[0] += s.value[k] * p[0];
    x[1] += s.value[k] * p[1];
  }
  return x;
}

std::array<double, 4> MacroMesh::jacobian(int32_t c, double xi, double eta) const {
  const CellShape s(npc_, xi, eta);
  std::array<double, 4> j{0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < s.n; ++k) {
    const double* p = &nodes_[2 * static_cast<std::size_t>(cell_node(c, k))];
    j[0] += s.d_xi[k] * p[0];
    j[1] += s.d_eta[k] * p[0];
    j[2] += s.d_xi[k] * p[1];
    j[3] += s.d_eta[k] * p[1];
  }
  return j;
}

// ------------------------------------------------------------------ forest
//
// A tree is its macro cell's reference square, cut into N x N max-level
// cells (N = 2^kMaxLevel).  A leaf (level L, i, j) covers the 2^(kMaxLevel-L)
// wide block of them whose lower-left cell is its anchor.

namespace {

constexpr int64_t kCells = int64_t{1} << Forest::kMaxLevel;

// Interleaves x into the even and y into the odd bits.
uint64_t morton(uint64_t x, uint64_t y) {
  auto spread = [](uint64_t v) {
    v &= 0xffffffffULL;
    v = (v | (v << 16)) & 0x0000ffff0000ffffULL;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffULL;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | (v << 2)) & 0x3333333333333333ULL;
    v = (v | (v << 1)) & 0x5555555555555555ULL;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

int32_t compact_even_bits(uint64_t v) {
  v &= 0x5555555555555555ULL;
  v = (v | (v >> 1)) & 0x3333333333333333ULL;
  v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0fULL;
  v = (v | (v >> 4)) & 0x00ff00ff00ff00ffULL;
  v = (v | (v >> 8)) & 0x0000ffff0000ffffULL;
  v = (v | (v >> 16)) & 0x00000000ffffffffULL;
  return static_cast<int32_t>(v);
}

int64_t block(const Leaf& l) { return int64_t{1} << (Forest::kMaxLevel - l.level); }

// Position of max-level cell (x, y) next to face f counted along the face
// counter-clockwise, and the cell of a tree just inside face f at such a
// position.
int64_t along_ccw(int f, int64_t x, int64_t y) {
  switch (f) {
    case 0: return x;
    case 1: return y;
    case 2: return kCells - 1 - x;
    default: return kCells - 1 - y;
  }
}

std::pair<int64_t, int64_t> inside_face(int f, int64_t u) {
  switch (f) {
    case 0: return {u, 0};
    case 1: return {kCells - 1, u};
    case 2: return {kCells - 1 - u, kCells - 1};
    default: return {0, kCells - 1 - u};
  }
}

// A lattice vertex (max-level lattice corner (X, Y) of a tree, 0..N) under
// an identity shared by every tree that has it: a macro node, a position
// along a macro edge in the edge's own direction, or a point inside one
// tree.
struct VertexKey {
  uint64_t hi, lo;
  bool operator==(const VertexKey& o) const { return hi == o.hi && lo == o.lo; }
};

struct VertexKeyHash {
  std::size_t operator()(const VertexKey& k) const {
    return static_cast<std::size_t>(k.hi * 0x9e3779b97f4a7c15ULL ^ (k.lo + 0x7f4a7c159e3779b9ULL + (k.hi << 7)));
  }
};

VertexKey vertex_key(const MacroMesh& m, int32_t tree, int64_t X, int64_t Y) {
  const bool x_edge = X == 0 || X == kCells, y_edge = Y == 0 || Y == kCells;
  if (x_edge && y_edge) {
    const int corner = Y == 0 ? (X == 0 ? 0 : 1) : (X == kCells ? 2 : 3);
    return {0, static_cast<uint64_t>(m.cell_node(tree, corner))};
  }
  if (x_edge || y_edge) {
    const int f = Y == 0 ? 0 : X == kCells ? 1 : Y == kCells ? 2 : 3;
    const int64_t ccw = f == 0 ? X : f == 1 ? Y : f == 2 ? kCells - X : kCells - Y;
    const int64_t pos = m.cell_edge_forward(tree, f) ? ccw : kCells - ccw;
    return {1ULL << 62 | static_cast<uint64_t>(m.cell_edge(tree, f)), static_cast<uint64_t>(pos)};
  }
  return {2ULL << 62 | static_cast<uint64_t>(tree), static_cast<uint64_t>(X) << 32 | static_cast<uint64_t>(Y)};
}

}  // namespace

Forest::Forest(std::shared_ptr<const MacroMesh> macro, std::vector<Leaf> leaves)
    : macro_(std::move(macro)), leaves_(std::move(leaves)) {
  code_.resize(leaves_.size());
  tree_first_.assign(static_cast<std::size_t>(macro_->n_cells()) + 1, 0);
  for (std::size_t l = 0; l < leaves_.size(); ++l) {
    const Leaf& lf = leaves_[l];
    code_[l] = morton(static_cast<uint64_t>(lf.i) * block(lf), static_cast<uint64_t>(lf.j) * block(lf));
    INSEM_REQUIRE(l == 0 || leaves_[l - 1].tree < lf.tree || (leaves_[l - 1].tree == lf.tree && code_[l - 1] < code_[l]),
                  "leaves must be in tree order and Morton order within a tree");
    ++tree_first_[static_cast<std::size_t>(lf.tree) + 1];
  }
  for (std::size_t t = 1; t < tree_first_.size(); ++t) tree_first_[t] += tree_first_[t - 1];
}

int32_t Forest::leaf_at(int32_t tree, int64_t x, int64_t y) const {
  if (x < 0 || y < 0 || x >= kCells || y >= kCells) return -1;
  const uint64_t c = morton(static_cast<uint64_t>(x), static_cast<uint64_t>(y));
  const auto first = code_.begin() + tree_first_[tree], last = code_.begin() + tree_first_[tree + 1];
  auto it = std::upper_bound(first, last, c);
  if (it == first) return -1;
  const int32_t l = static_cast<int32_t>((it - 1) - code_.begin());
  const Leaf& lf = leaves_[l];
  const int64_t b = block(lf);
  return x >= lf.i * b && x < (lf.i + 1) * b && y >= lf.j * b && y < (lf.j + 1) * b ? l : -1;
}

Forest Forest::uniform(std::shared_ptr<const MacroMesh> macro, int level) {
  INSEM_REQUIRE(macro, "a forest needs a macro mesh");
  INSEM_REQUIRE(level >= 0 && level <= kMaxLevel, "uniform level " << level << " is outside 0.." << kMaxLevel);
  const uint64_t per_tree = uint64_t{1} << (2 * level);
  INSEM_REQUIRE(per_tree * static_cast<uint64_t>(macro->n_cells()) < (uint64_t{1} << 31),
                "level " << level << " on " << macro->n_cells() << " cells makes too many leaves");
  std::vector<Leaf> leaves;
  leaves.reserve(static_cast<std::size_t>(per_tree * macro->n_cells()));
  for (int32_t t = 0; t < macro->n_cells(); ++t)
    for (uint64_t m = 0; m < per_tree; ++m) leaves.push_back({t, level, compact_even_bits(m), compact_even_bits(m >> 1)});
  return Forest(std::move(macro), std::move(leaves));
}

int Forest::max_level() const {
  int top = 0;
  for (const Leaf& l : leaves_) top = std::max(top, static_cast<int>(l.level));
  return top;
}

std::array<double, 4> Forest::leaf_box(int32_t l) const {
  const Leaf& lf = leaves_[l];
  const double h = 2.0 / static_cast<double>(int64_t{1} << lf.level);
  return {-1.0 + lf.i * h, -1.0 + (lf.i + 1) * h, -1.0 + lf.j * h, -1.0 + (lf.j + 1) * h};
}

// The leaf across face f: the one covering the max-level cell just outside
// the face at (or just past) its midpoint, carried across a macro edge in
// the edge's own direction.
Neighbor Forest::neighbor(int32_t l, int f) const {
  const Leaf& lf = leaves_[l];
  const int64_t b = block(lf), x0 = lf.i * b, y0 = lf.j * b, half = b / 2;
  int64_t x, y;
  switch (f) {
    case 0: x = x0 + half, y = y0 - 1; break;
    case 1: x = x0 + b, y = y0 + half; break;
    case 2: x = x0 + half, y = y0 + b; break;
    default: x = x0 - 1, y = y0 + half; break;
  }
  Neighbor out;
  int32_t tree = lf.tree;
  out.face = (f + 2) % 4;
  if (x < 0 || y < 0 || x >= kCells || y >= kCells) {
    const MacroMesh& m = *macro_;
    const int32_t e = m.cell_edge(lf.tree, f);
    if (m.is_boundary_edge(e)) return Neighbor{};
    const auto cells = m.edge_cells(e);
    const int32_t other = cells[0] == lf.tree ? cells[1] : cells[0];
    int of = 0;
    while (m.cell_edge(other, of) != e) ++of;
    // along the face counter-clockwise, in the edge's direction, in the other cell's
    const int64_t u = along_ccw(f, std::clamp<int64_t>(x, 0, kCells - 1), std::clamp<int64_t>(y, 0, kCells - 1));
    const int64_t on_edge = m.cell_edge_forward(lf.tree, f) ? u : kCells - 1 - u;
    const int64_t u2 = m.cell_edge_forward(other, of) ? on_edge : kCells - 1 - on_edge;
    std::tie(x, y) = inside_face(of, u2);
    tree = other;
    out.face = of;
  }
  out.leaf = leaf_at(tree, x, y);
  INSEM_REQUIRE(out.leaf >= 0, "no leaf covers the position across face " << f << " of leaf " << l);
  out.level_diff = leaves_[out.leaf].level - lf.level;
  return out;
}

// Coordinate t in [-1, 1] along face f of leaf `from` (in the leaf's i
// direction on faces 0 and 2, j on faces 1 and 3) as the coordinate along
// face to_face of leaf `to`, which shares that stretch of boundary.
double Forest::face_coordinate(int32_t from, int f, double t, int32_t to, int to_face) const {
  // the face's stretch of its tree's side, as fractions [lo, hi] of the side
  auto stretch = [this](int32_t l, int face) {
    const Leaf& lf = leaves_[l];
    const double cells = static_cast<double>(int64_t{1} << lf.level);
    const int32_t k = face == 0 || face == 2 ? lf.i : lf.j;
    return std::array<double, 2>{k / cells, (k + 1) / cells};
  };
  const auto a = stretch(from, f);
  double u = a[0] + 0.5 * (t + 1.0) * (a[1] - a[0]);
  const Leaf& lf = leaves_[from];
  const Leaf& lt = leaves_[to];
  if (lf.tree != lt.tree) {
    const MacroMesh& m = *macro_;
    // i/j fraction -> counter-clockwise -> edge direction -> the other side
    const double ccw = f == 0 || f == 1 ? u : 1.0 - u;
    const double on_edge = m.cell_edge_forward(lf.tree, f) ? ccw : 1.0 - ccw;
    const double ccw2 = m.cell_edge_forward(lt.tree, to_face) ? on_edge : 1.0 - on_edge;
    u = to_face == 0 || to_face == 1 ? ccw2 : 1.0 - ccw2;
  }
  const auto b = stretch(to, to_face);
  return -1.0 + 2.0 * (u - b[0]) / (b[1] - b[0]);
}

// Marked leaves are replaced in place by their four children, in Morton
// order.
Forest Forest::refine(const std::vector<char>& mark) const {
  INSEM_REQUIRE(mark.size() == leaves_.size(), mark.size() << " refinement marks for " << leaves_.size() << " leaves");
  std::vector<Leaf> out;
  out.reserve(leaves_.size() + 3 * static_cast<std::size_t>(std::count(mark.begin(), mark.end(), 1)));
  for (std::size_t l = 0; l < leaves_.size(); ++l) {
    const Leaf& lf = leaves_[l];
    if (!mark[l]) {
      out.push_back(lf);
      continue;
    }
    INSEM_REQUIRE(lf.level < kMaxLevel, "leaf " << l << " is already at the finest level " << kMaxLevel);
    for (int c = 0; c < 4; ++c) out.push_back({lf.tree, lf.level + 1, 2 * lf.i + (c & 1), 2 * lf.j + (c >> 1)});
  }
  return Forest(macro_, std::move(out));
}

// Sweeps until no face joins leaves two or more levels apart (and, across
// the tree pairs strict_pair names, none joins different levels at all):
// each sweep refines every coarser leaf that breaks the rule.
Forest Forest::balance(const std::function<bool(int32_t, int32_t)>& strict_pair) const {
  Forest cur = *this;
  for (;;) {
    std::vector<char> mark(cur.leaves_.size(), 0);
    bool changed = false;
    for (int32_t l = 0; l < cur.n_leaves(); ++l)
      for (int f = 0; f < 4; ++f) {
        const Neighbor nb = cur.neighbor(l, f);
        if (nb.leaf < 0 || nb.level_diff >= 0) continue;
        const int32_t a = cur.leaves_[l].tree, b = cur.leaves_[nb.leaf].tree;
        if (nb.level_diff <= -2 || (a != b && strict_pair && strict_pair(a, b))) {
          mark[nb.leaf] = 1;
          changed = true;
        }
      }
    if (!changed) return cur;
    cur = cur.refine(mark);
  }
}

// Lattice vertices numbered in order of first appearance over the leaves'
// corners (counter-clockwise from the lower left).
Corners Forest::corners() const {
  Corners out;
  out.corner.resize(4 * leaves_.size());
  std::unordered_map<VertexKey, int32_t, VertexKeyHash> index;
  index.reserve(leaves_.size() + static_cast<std::size_t>(macro_->n_nodes()));
  constexpr int dx[4] = {0, 1, 1, 0}, dy[4] = {0, 0, 1, 1};
  for (std::size_t l = 0; l < leaves_.size(); ++l) {
    const Leaf& lf = leaves_[l];
    const int64_t b = block(lf);
    for (int k = 0; k < 4; ++k) {
      const VertexKey key = vertex_key(*macro_, lf.tree, (lf.i + dx[k]) * b, (lf.j + dy[k]) * b);
      const auto ins = index.emplace(key, out.n_vertices);
      if (ins.second) ++out.n_vertices;
      out.corner[4 * l + k] = ins.first->second;
    }
  }
  return out;
}

}  // namespace insem_amd
