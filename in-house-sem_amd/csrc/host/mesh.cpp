#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

#include "error.hpp"

namespace insem_amd {
namespace {

// Lattice position (0 = -1, 1 = 0, 2 = +1) of the 9 biquadratic nodes:
// corners counter-clockwise, face midpoints 0..3, centre.
constexpr int kQ9a[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
constexpr int kQ9b[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
constexpr double kQ4xi[4] = {-1.0, 1.0, 1.0, -1.0};
constexpr double kQ4eta[4] = {-1.0, -1.0, 1.0, 1.0};

// Shape functions and reference derivatives (cpp/src/mesh/macro_mesh.cpp:19-50).
int shape_functions(int npc, double xi, double eta, double* n, double* nx, double* ne) {
  if (npc == 4) {
    for (int k = 0; k < 4; ++k) {
      const double a = 1.0 + kQ4xi[k] * xi, b = 1.0 + kQ4eta[k] * eta;
      n[k] = 0.25 * a * b;
      nx[k] = 0.25 * kQ4xi[k] * b;
      ne[k] = 0.25 * a * kQ4eta[k];
    }
    return 4;
  }
  auto quad = [](double t, double* l, double* dl) {
    l[0] = 0.5 * t * (t - 1.0);
    l[1] = 1.0 - t * t;
    l[2] = 0.5 * t * (t + 1.0);
    dl[0] = t - 0.5;
    dl[1] = -2.0 * t;
    dl[2] = t + 0.5;
  };
  double lx[3], dlx[3], ly[3], dly[3];
  quad(xi, lx, dlx);
  quad(eta, ly, dly);
  for (int k = 0; k < 9; ++k) {
    n[k] = lx[kQ9a[k]] * ly[kQ9b[k]];
    nx[k] = dlx[kQ9a[k]] * ly[kQ9b[k]];
    ne[k] = lx[kQ9a[k]] * dly[kQ9b[k]];
  }
  return 9;
}

}  // namespace

MacroMesh::MacroMesh(int nodes_per_cell, bool axisymmetric, std::vector<double> nodes, std::vector<int32_t> cells,
                     std::vector<int> material, std::vector<BoundaryFace> boundary)
    : npc_(nodes_per_cell),
      axisym_(axisymmetric),
      nodes_(std::move(nodes)),
      cells_(std::move(cells)),
      material_(std::move(material)),
      boundary_(std::move(boundary)) {
  INSEM_REQUIRE(npc_ == 4 || npc_ == 9, "cells must have 4 or 9 nodes, got " << npc_);
  INSEM_REQUIRE(nodes_.size() % 2 == 0, "node coordinate array must hold (x, z) pairs");
  INSEM_REQUIRE(cells_.size() == material_.size() * static_cast<std::size_t>(npc_),
                "cell array has " << cells_.size() << " entries, expected " << material_.size() * npc_);
  for (int32_t v : cells_) INSEM_REQUIRE(v >= 0 && v < n_nodes(), "cell references node " << v << " out of range");
  if (axisym_) rotate_axis_cells();
  build_edges();
  check();
}

bool MacroMesh::is_axis_face(int32_t c, int f) const {
  if (!axisym_) return false;
  double ext = 0.0;
  for (int k = 0; k < 4; ++k) {
    const int32_t v = cell_node(c, k);
    ext = std::max({ext, std::fabs(nodes_[2 * v]), std::fabs(nodes_[2 * v + 1])});
  }
  const double tol = 1e-10 * std::max(ext, 1e-300);
  return std::fabs(nodes_[2 * cell_node(c, kFaceCorner[f][0])]) <= tol &&
         std::fabs(nodes_[2 * cell_node(c, kFaceCorner[f][1])]) <= tol;
}

// Cells with a face on the axis are renumbered so that face is face 3
// (cpp/src/mesh/macro_mesh.cpp:105-129): new corner i = old corner (i + k) mod 4.
void MacroMesh::rotate_axis_cells() {
  for (int32_t c = 0; c < n_cells(); ++c) {
    int af = -1;
    for (int f = 0; f < 4; ++f) {
      if (!is_axis_face(c, f)) continue;
      INSEM_REQUIRE(af < 0, "cell " << c << " has more than one face on the axis");
      af = f;
    }
    if (af < 0 || af == 3) continue;
    const int k = (af + 1) % 4;
    int32_t* cn = &cells_[static_cast<std::size_t>(c) * npc_];
    const std::vector<int32_t> old(cn, cn + npc_);
    for (int i = 0; i < 4; ++i) cn[i] = old[(i + k) % 4];
    if (npc_ == 9)
      for (int i = 0; i < 4; ++i) cn[4 + i] = old[4 + (i + k) % 4];
    for (auto& bf : boundary_)
      if (bf.cell == c) bf.face = (bf.face - k + 4) % 4;
  }
}

// Edge ids in order of first appearance; the canonical direction runs from
// the lower to the higher node index (cpp/src/mesh/macro_mesh.cpp:131-171).
void MacroMesh::build_edges() {
  const int32_t nc = n_cells();
  cell_edges_.assign(4 * static_cast<std::size_t>(nc), -1);
  edge_fwd_.assign(4 * static_cast<std::size_t>(nc), 0);
  edge_cells_.clear();
  std::unordered_map<uint64_t, int32_t> ids;
  ids.reserve(4 * static_cast<std::size_t>(nc));
  for (int32_t c = 0; c < nc; ++c) {
    for (int f = 0; f < 4; ++f) {
      const int32_t a = cell_node(c, kFaceCorner[f][0]), b = cell_node(c, kFaceCorner[f][1]);
      INSEM_REQUIRE(a != b, "cell " << c << " has a degenerate face " << f);
      const uint64_t key = (static_cast<uint64_t>(std::min(a, b)) << 32) | static_cast<uint32_t>(std::max(a, b));
      auto ins = ids.emplace(key, n_edges());
      const int32_t e = ins.first->second;
      if (ins.second) {
        edge_cells_.push_back(c);
        edge_cells_.push_back(-1);
      } else {
        INSEM_REQUIRE(edge_cells_[2 * e + 1] < 0, "edge (" << a << ", " << b << ") is shared by more than two cells");
        const int32_t o = edge_cells_[2 * e];
        int of = 0;
        while (cell_edges_[4 * static_cast<std::size_t>(o) + of] != e) ++of;
        INSEM_REQUIRE((a < b) != (edge_fwd_[4 * static_cast<std::size_t>(o) + of] != 0),
                      "cells " << o << " and " << c << " have inconsistent orientation");
        edge_cells_[2 * e + 1] = c;
      }
      cell_edges_[4 * static_cast<std::size_t>(c) + f] = e;
      edge_fwd_[4 * static_cast<std::size_t>(c) + f] = a < b ? 1 : 0;
    }
  }
}

void MacroMesh::check() const {
  for (int32_t c = 0; c < n_cells(); ++c)
    for (double xi : {-1.0, 0.0, 1.0})
      for (double eta : {-1.0, 0.0, 1.0}) {
        const auto j = jacobian(c, xi, eta);
        const double det = j[0] * j[3] - j[1] * j[2];
        INSEM_REQUIRE(det > 0.0, "cell " << c << " has a non-positive Jacobian (" << det
                                         << "); cells must be counter-clockwise and non-degenerate");
      }
  if (axisym_) {
    double ext = 0.0;
    for (double v : nodes_) ext = std::max(ext, std::fabs(v));
    for (int32_t i = 0; i < n_nodes(); ++i)
      INSEM_REQUIRE(nodes_[2 * i] >= -1e-12 * ext, "axisymmetric mesh has node " << i << " at negative r");
  }
  for (const auto& bf : boundary_) {
    INSEM_REQUIRE(bf.cell >= 0 && bf.cell < n_cells(), "boundary face references cell " << bf.cell);
    INSEM_REQUIRE(bf.face >= 0 && bf.face < 4, "boundary face index " << bf.face << " not in [0, 4)");
    INSEM_REQUIRE(is_boundary_edge(cell_edge(bf.cell, bf.face)),
                  "face " << bf.face << " of cell " << bf.cell << " is tagged as boundary but is shared");
  }
}

std::array<double, 2> MacroMesh::map(int32_t c, double xi, double eta) const {
  double n[9], nx[9], ne[9];
  const int m = shape_functions(npc_, xi, eta, n, nx, ne);
  std::array<double, 2> x{0.0, 0.0};
  for (int k = 0; k < m; ++k) {
    const int32_t v = cell_node(c, k);
    x[0] += n[k] * nodes_[2 * v];
    x[1] += n[k] * nodes_[2 * v + 1];
  }
  return x;
}

std::array<double, 4> MacroMesh::jacobian(int32_t c, double xi, double eta) const {
  double n[9], nx[9], ne[9];
  const int m = shape_functions(npc_, xi, eta, n, nx, ne);
  std::array<double, 4> j{0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < m; ++k) {
    const int32_t v = cell_node(c, k);
    j[0] += nx[k] * nodes_[2 * v];
    j[1] += ne[k] * nodes_[2 * v];
    j[2] += nx[k] * nodes_[2 * v + 1];
    j[3] += ne[k] * nodes_[2 * v + 1];
  }
  return j;
}

// ------------------------------------------------------------------ forest

namespace {

// Identity of a corner of the 2^kMaxLevel lattice of a tree that is the same
// from every tree sharing it (cpp/src/mesh/forest.cpp:15-62).
struct PointId {
  int kind;  // 0 macro node, 1 macro edge, 2 tree interior
  int64_t id, a, b;
  bool operator==(const PointId& o) const { return kind == o.kind && id == o.id && a == o.a && b == o.b; }
};

struct PointIdHash {
  std::size_t operator()(const PointId& p) const {
    uint64_t h = static_cast<uint64_t>(p.kind);
    for (int64_t v : {p.id, p.a, p.b}) h ^= static_cast<uint64_t>(v) + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
    return static_cast<std::size_t>(h);
  }
};

PointId point_id(const MacroMesh& m, int32_t cell, int64_t a, int64_t b) {
  constexpr int64_t s = int64_t{1} << Forest::kMaxLevel;
  const bool l = a == 0, r = a == s, bot = b == 0, top = b == s;
  if ((l || r) && (bot || top)) {
    const int corner = bot ? (l ? 0 : 1) : (r ? 2 : 3);
    return {0, m.cell_node(cell, corner), 0, 0};
  }
  if (bot || r || top || l) {
    int f;
    int64_t t;
    if (bot) {
      f = 0;
      t = a;
    } else if (r) {
      f = 1;
      t = b;
    } else if (top) {
      f = 2;
      t = s - a;
    } else {
      f = 3;
      t = s - b;
    }
    return {1, m.cell_edge(cell, f), m.cell_edge_forward(cell, f) ? t : s - t, 0};
  }
  return {2, cell, a, b};
}

// Even bits of a Morton code.
int32_t deinterleave(uint64_t m) {
  uint64_t x = m & 0x5555555555555555ULL;
  x = (x | (x >> 1)) & 0x3333333333333333ULL;
  x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0fULL;
  x = (x | (x >> 4)) & 0x00ff00ff00ff00ffULL;
  x = (x | (x >> 8)) & 0x0000ffff0000ffffULL;
  x = (x | (x >> 16)) & 0x00000000ffffffffULL;
  return static_cast<int32_t>(x);
}

}  // namespace

Forest::Forest(std::shared_ptr<const MacroMesh> macro, std::vector<Leaf> leaves)
    : macro_(std::move(macro)), leaves_(std::move(leaves)) {
  lookup_.reserve(leaves_.size());
  for (std::size_t l = 0; l < leaves_.size(); ++l) {
    const Leaf& lf = leaves_[l];
    lookup_[key(lf.tree, lf.level, lf.i, lf.j)] = static_cast<int32_t>(l);
  }
}

std::pair<uint64_t, uint64_t> Forest::key(int32_t tree, int level, int64_t i, int64_t j) {
  return {(static_cast<uint64_t>(tree) << 8) | static_cast<uint64_t>(level),
          (static_cast<uint64_t>(i) << 32) | static_cast<uint64_t>(j)};
}

Forest Forest::uniform(std::shared_ptr<const MacroMesh> macro, int level) {
  INSEM_REQUIRE(macro, "forest needs a macro-mesh");
  INSEM_REQUIRE(level >= 0 && level <= kMaxLevel, "refinement level " << level << " not in [0, " << kMaxLevel << "]");
  const uint64_t per = uint64_t{1} << (2 * level);
  const uint64_t total = per * static_cast<uint64_t>(macro->n_cells());
  INSEM_REQUIRE(total < (uint64_t{1} << 31), "uniform refinement to level " << level << " gives too many leaves");
  std::vector<Leaf> leaves;
  leaves.reserve(static_cast<std::size_t>(total));
  for (int32_t t = 0; t < macro->n_cells(); ++t)
    for (uint64_t m = 0; m < per; ++m) leaves.push_back({t, level, deinterleave(m), deinterleave(m >> 1)});
  return Forest(std::move(macro), std::move(leaves));
}

int Forest::max_level() const {
  int m = 0;
  for (const Leaf& l : leaves_) m = std::max(m, static_cast<int>(l.level));
  return m;
}

std::array<double, 4> Forest::leaf_box(int32_t l) const {
  const Leaf& lf = leaves_[l];
  const double h = 2.0 / static_cast<double>(int64_t{1} << lf.level);
  return {-1.0 + lf.i * h, -1.0 + (lf.i + 1) * h, -1.0 + lf.j * h, -1.0 + (lf.j + 1) * h};
}

int32_t Forest::find_leaf(int32_t tree, int64_t a, int64_t b) const {
  constexpr int64_t h = int64_t{1} << (kMaxLevel + 1);
  if (a < 0 || b < 0 || a >= h || b >= h) return -1;
  for (int level = 0; level <= kMaxLevel; ++level) {
    const int sh = kMaxLevel + 1 - level;
    const auto it = lookup_.find(key(tree, level, a >> sh, b >> sh));
    if (it != lookup_.end()) return it->second;
  }
  return -1;
}

// Probes the half-lattice point just outside the face midpoint; across a
// macro edge the position is carried over in canonical edge direction
// (cpp/src/mesh/forest.cpp:134-176).
Neighbor Forest::neighbor(int32_t l, int f) const {
  constexpr int64_t h = int64_t{1} << (kMaxLevel + 1);
  const Leaf& lf = leaves_[l];
  const int64_t size = int64_t{1} << (kMaxLevel + 1 - lf.level);
  const int64_t a0 = lf.i * size, b0 = lf.j * size;
  int64_t a, b;
  switch (f) {
    case 0: a = a0 + size / 2; b = b0 - 1; break;
    case 1: a = a0 + size; b = b0 + size / 2; break;
    case 2: a = a0 + size / 2; b = b0 + size; break;
    default: a = a0 - 1; b = b0 + size / 2; break;
  }
  Neighbor out;
  int32_t tree = lf.tree;
  int nf = (f + 2) % 4;
  if (a < 0 || b < 0 || a >= h || b >= h) {
    const MacroMesh& m = *macro_;
    const int32_t e = m.cell_edge(lf.tree, f);
    const auto cells = m.edge_cells(e);
    if (cells[1] < 0) return out;
    const int32_t other = cells[0] == lf.tree ? cells[1] : cells[0];
    int of = 0;
    while (m.cell_edge(other, of) != e) ++of;
    const int64_t sl = f == 0 ? a : f == 1 ? b : f == 2 ? h - a : h - b;
    const int64_t canon = m.cell_edge_forward(lf.tree, f) ? sl : h - sl;
    const int64_t s2 = m.cell_edge_forward(other, of) ? canon : h - canon;
    switch (of) {
      case 0: a = s2; b = 0; break;
      case 1: a = h - 1; b = s2; break;
      case 2: a = h - s2; b = h - 1; break;
      default: a = 0; b = h - s2; break;
    }
    tree = other;
    nf = of;
  }
  out.leaf = find_leaf(tree, a, b);
  INSEM_REQUIRE(out.leaf >= 0, "forest has no leaf at a face-neighbour position");
  out.face = nf;
  out.level_diff = leaves_[out.leaf].level - lf.level;
  return out;
}

double Forest::face_coordinate(int32_t from, int f, double t, int32_t to, int to_face) const {
  auto range = [this](int32_t l, int face) {
    const Leaf& lf = leaves_[l];
    const double n = static_cast<double>(int64_t{1} << lf.level);
    const int64_t idx = (face == 0 || face == 2) ? lf.i : lf.j;
    return std::array<double, 2>{idx / n, (idx + 1) / n};
  };
  const auto r = range(from, f);
  const double u = r[0] + 0.5 * (t + 1.0) * (r[1] - r[0]);
  double u2 = u;
  const Leaf& a = leaves_[from];
  const Leaf& b = leaves_[to];
  if (a.tree != b.tree) {
    const MacroMesh& m = *macro_;
    const double s = (f == 0 || f == 1) ? u : 1.0 - u;
    const double c = m.cell_edge_forward(a.tree, f) ? s : 1.0 - s;
    const double s2 = m.cell_edge_forward(b.tree, to_face) ? c : 1.0 - c;
    u2 = (to_face == 0 || to_face == 1) ? s2 : 1.0 - s2;
  }
  const auto r2 = range(to, to_face);
  return -1.0 + 2.0 * (u2 - r2[0]) / (r2[1] - r2[0]);
}

// Children in Morton order replace their parent in place.
Forest Forest::refine(const std::vector<char>& mark) const {
  INSEM_REQUIRE(mark.size() == leaves_.size(), "refinement marks do not match the leaves");
  std::vector<Leaf> out;
  out.reserve(leaves_.size() + 3 * static_cast<std::size_t>(std::count(mark.begin(), mark.end(), 1)));
  for (std::size_t l = 0; l < leaves_.size(); ++l) {
    const Leaf& lf = leaves_[l];
    if (!mark[l]) {
      out.push_back(lf);
      continue;
    }
    INSEM_REQUIRE(lf.level < kMaxLevel, "cannot refine beyond level " << kMaxLevel);
    for (int c = 0; c < 4; ++c) out.push_back({lf.tree, lf.level + 1, 2 * lf.i + (c & 1), 2 * lf.j + (c >> 1)});
  }
  return Forest(macro_, std::move(out));
}

// Repeated sweeps: a coarser neighbour more than one level up, or any
// coarser neighbour across a strict tree pair, is marked and refined.
Forest Forest::balance(const std::function<bool(int32_t, int32_t)>& strict_pair) const {
  Forest cur = *this;
  for (;;) {
    std::vector<char> mark(cur.leaves_.size(), 0);
    bool any = false;
    for (int32_t l = 0; l < cur.n_leaves(); ++l) {
      const Leaf& lf = cur.leaves_[l];
      for (int f = 0; f < 4; ++f) {
        const Neighbor nb = cur.neighbor(l, f);
        if (nb.leaf < 0 || nb.level_diff >= 0) continue;
        const int32_t ot = cur.leaves_[nb.leaf].tree;
        const bool strict = strict_pair && ot != lf.tree && strict_pair(lf.tree, ot);
        if (nb.level_diff < -1 || strict) {
          mark[nb.leaf] = 1;
          any = true;
        }
      }
    }
    if (!any) return cur;
    cur = cur.refine(mark);
  }
}

Corners Forest::corners() const {
  Corners out;
  out.corner.resize(4 * leaves_.size());
  std::unordered_map<PointId, int32_t, PointIdHash> ids;
  ids.reserve(leaves_.size() + static_cast<std::size_t>(macro_->n_nodes()));
  constexpr int di[4] = {0, 1, 1, 0};
  constexpr int dj[4] = {0, 0, 1, 1};
  for (std::size_t l = 0; l < leaves_.size(); ++l) {
    const Leaf& lf = leaves_[l];
    const int64_t step = int64_t{1} << (kMaxLevel - lf.level);
    for (int k = 0; k < 4; ++k) {
      const PointId p = point_id(*macro_, lf.tree, (lf.i + di[k]) * step, (lf.j + dj[k]) * step);
      auto ins = ids.emplace(p, out.n_vertices);
      if (ins.second) ++out.n_vertices;
      out.corner[4 * l + k] = ins.first->second;
    }
  }
  return out;
}

}  // namespace insem_amd
