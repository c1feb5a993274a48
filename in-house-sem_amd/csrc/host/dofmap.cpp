#include "dofmap.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <unordered_map>

#include "basis.hpp"
#include "error.hpp"

namespace insem_amd {
namespace {

constexpr int32_t kUnresolved = std::numeric_limits<int32_t>::min();

// Local index of point m along the axis of face f.
int face_local(int f, int m, int n, int ngll) {
  switch (f) {
    case 0: return m;
    case 1: return m * ngll + n;
    case 2: return n * ngll + m;
    default: return m * ngll;
  }
}

}  // namespace

int32_t DofMap::global_index(int32_t e, int k) const {
  const int32_t m = l2g[static_cast<std::size_t>(e) * npe + k];
  INSEM_REQUIRE(m >= 0, "local point " << k << " of element " << e << " is constrained by a hanging face");
  return m;
}

std::vector<std::pair<int32_t, double>> DofMap::slot_entries(int32_t e, int k) const {
  const int32_t m = l2g[static_cast<std::size_t>(e) * npe + k];
  if (m >= 0) return {{m, 1.0}};
  std::vector<std::pair<int32_t, double>> out;
  const int32_t r = -1 - m;
  for (int32_t q = row_offsets[r]; q < row_offsets[r + 1]; ++q) out.push_back({row_dofs[q], row_weights[q]});
  return out;
}

std::vector<int32_t> DofMap::multiplicity() const {
  std::vector<int32_t> m(n_global);
  for (int32_t g = 0; g < n_global; ++g) m[g] = t_offsets[g + 1] - t_offsets[g];
  return m;
}

DofMap DofMap::build(const Forest& forest, int ngll, const std::vector<int32_t>& leaves) {
  INSEM_REQUIRE(ngll >= 2, "need at least 2 GLL points per direction, got " << ngll);
  DofMap d;
  d.ngll = ngll;
  d.npe = ngll * ngll;
  d.n_elements = static_cast<int32_t>(leaves.size());
  const int n = ngll - 1;
  const std::size_t npe = static_cast<std::size_t>(d.npe);
  const std::size_t nslots = static_cast<std::size_t>(d.n_elements) * npe;
  INSEM_REQUIRE(nslots < static_cast<std::size_t>(std::numeric_limits<int32_t>::max()),
                "too many element-local points (" << nslots << ") for 32-bit indices");
  std::vector<int32_t> elem_of_leaf(forest.n_leaves(), -1);
  for (int32_t e = 0; e < d.n_elements; ++e) {
    INSEM_REQUIRE(leaves[e] >= 0 && leaves[e] < forest.n_leaves(), "leaf index " << leaves[e] << " out of range");
    elem_of_leaf[leaves[e]] = e;
  }
  const Corners cv = forest.corners();

  // 1. Fine faces that hang on a coarse face of the same set, and the
  //    vertices in the middle of such coarse faces.
  struct Hang {
    int32_t coarse = -1;
    int face = -1;
  };
  std::vector<std::array<Hang, 4>> hang(d.n_elements);
  std::vector<char> mid_vertex(cv.n_vertices, 0);
  for (int32_t e = 0; e < d.n_elements; ++e) {
    const int32_t leaf = leaves[e];
    for (int f = 0; f < 4; ++f) {
      const Neighbor nb = forest.neighbor(leaf, f);
      if (nb.leaf < 0) continue;
      const int32_t ce = elem_of_leaf[nb.leaf];
      if (ce < 0) {
        INSEM_REQUIRE(nb.level_diff == 0, "leaves of different domains meet at different refinement levels");
        continue;
      }
      if (nb.level_diff >= 0) continue;
      INSEM_REQUIRE(nb.level_diff == -1, "forest is not 2:1 balanced (level jump " << -nb.level_diff << ")");
      hang[e][f] = {ce, nb.face};
      const int32_t* cc = &cv.corner[4 * static_cast<std::size_t>(nb.leaf)];
      const int32_t c0 = cc[kFaceCorner[nb.face][0]], c1 = cc[kFaceCorner[nb.face][1]];
      for (int s = 0; s < 2; ++s) {
        const int32_t v = cv.corner[4 * static_cast<std::size_t>(leaf) + kFaceCorner[f][s]];
        if (v != c0 && v != c1) mid_vertex[v] = 1;
      }
    }
  }

  // 2. First-touch numbering of the unconstrained points.
  std::vector<int32_t> vid(cv.n_vertices, -1);
  struct EdgeUse {
    int32_t base;
    int uses;
  };
  std::unordered_map<uint64_t, EdgeUse> edges;
  edges.reserve(2 * static_cast<std::size_t>(d.n_elements) + 16);
  d.l2g.assign(nslots, -1);
  int32_t next = 0;
  for (int32_t e = 0; e < d.n_elements; ++e) {
    const int32_t* v = &cv.corner[4 * static_cast<std::size_t>(leaves[e])];
    int32_t base[4];
    bool fwd[4];
    for (int f = 0; f < 4; ++f) {
      const int32_t a = v[kFaceCorner[f][0]], b = v[kFaceCorner[f][1]];
      fwd[f] = a < b;
      if (hang[e][f].coarse >= 0) {
        base[f] = -1;
        continue;
      }
      const uint64_t key = (static_cast<uint64_t>(std::min(a, b)) << 32) | static_cast<uint32_t>(std::max(a, b));
      auto ins = edges.emplace(key, EdgeUse{next, 0});
      if (ins.second) next += n - 1;
      ++ins.first->second.uses;
      INSEM_REQUIRE(ins.first->second.uses <= 2, "leaf edge shared by more than two elements");
      base[f] = ins.first->second.base;
    }
    int32_t* map = &d.l2g[static_cast<std::size_t>(e) * npe];
    for (int j = 0; j <= n; ++j) {
      for (int i = 0; i <= n; ++i) {
        const bool ie = i == 0 || i == n, je = j == 0 || j == n;
        int32_t g;
        if (ie && je) {
          const int corner = j == 0 ? (i == 0 ? 0 : 1) : (i == n ? 2 : 3);
          if (mid_vertex[v[corner]]) {
            g = kUnresolved;
          } else {
            int32_t& id = vid[v[corner]];
            if (id < 0) id = next++;
            g = id;
          }
        } else if (ie || je) {
          int f, p;
          if (j == 0) {
            f = 0;
            p = i;
          } else if (i == n) {
            f = 1;
            p = j;
          } else if (j == n) {
            f = 2;
            p = n - i;
          } else {
            f = 3;
            p = n - j;
          }
          g = base[f] < 0 ? kUnresolved : base[f] + ((fwd[f] ? p : n - p) - 1);
        } else {
          g = next++;
        }
        map[j * ngll + i] = g;
      }
    }
  }
  d.n_global = next;

  // 3. Constraint rows: value of a hanging slot = coarse face interpolant at
  //    its position, recursively through coarse slots that are constrained.
  const Lagrange1D gll_b(gll_rule(ngll).points);
  const Lagrange1D glj_b(glj01_rule(ngll).points);
  auto face_basis = [&](int32_t e, int f) -> const Lagrange1D& {
    return (forest.is_axial(leaves[e]) && (f == 0 || f == 2)) ? glj_b : gll_b;
  };
  using Row = std::vector<std::pair<int32_t, double>>;
  std::map<std::size_t, Row> memo;
  std::function<Row(std::size_t)> resolve = [&](std::size_t slot) -> Row {
    const int32_t m = d.l2g[slot];
    if (m >= 0) return {{m, 1.0}};
    if (m != kUnresolved) {
      const int32_t r = -1 - m;
      Row row;
      for (int32_t q = d.row_offsets[r]; q < d.row_offsets[r + 1]; ++q) row.push_back({d.row_dofs[q], d.row_weights[q]});
      return row;
    }
    const auto hit = memo.find(slot);
    if (hit != memo.end()) return hit->second;
    const int32_t e = static_cast<int32_t>(slot / npe);
    const int k = static_cast<int>(slot % npe);
    const int i = k % ngll, j = k / ngll;
    int f = -1;
    for (int c = 0; c < 4 && f < 0; ++c) {
      if (hang[e][c].coarse < 0) continue;
      if ((c == 0 && j == 0) || (c == 1 && i == n) || (c == 2 && j == n) || (c == 3 && i == 0)) f = c;
    }
    INSEM_REQUIRE(f >= 0, "constrained point is not on a hanging face");
    const Hang h = hang[e][f];
    const double t = face_basis(e, f).nodes()[(f == 0 || f == 2) ? i : j];
    const double tc = forest.face_coordinate(leaves[e], f, t, leaves[h.coarse], h.face);
    const auto w = face_basis(h.coarse, h.face).values(tc);
    std::map<int32_t, double> acc;
    for (int q = 0; q <= n; ++q) {
      if (std::fabs(w[q]) < 1e-14) continue;
      const std::size_t cs = static_cast<std::size_t>(h.coarse) * npe + face_local(h.face, q, n, ngll);
      for (const auto& [g, wg] : resolve(cs)) acc[g] += w[q] * wg;
    }
    Row row(acc.begin(), acc.end());
    memo[slot] = row;
    return row;
  };
  for (std::size_t s = 0; s < nslots; ++s) {
    if (d.l2g[s] != kUnresolved) continue;
    const Row row = resolve(s);
    if (row.size() == 1 && std::fabs(row[0].second - 1.0) < 1e-12) {
      d.l2g[s] = row[0].first;
      continue;
    }
    d.l2g[s] = -1 - d.n_rows();
    for (const auto& [g, w] : row) {
      d.row_dofs.push_back(g);
      d.row_weights.push_back(w);
    }
    d.row_offsets.push_back(static_cast<int32_t>(d.row_dofs.size()));
  }

  // 4. Transpose CSR in ascending slot order.
  auto visit = [&](auto&& fn) {
    for (std::size_t s = 0; s < nslots; ++s) {
      const int32_t m = d.l2g[s];
      if (m >= 0) {
        fn(s, m, 1.0);
      } else {
        const int32_t r = -1 - m;
        for (int32_t q = d.row_offsets[r]; q < d.row_offsets[r + 1]; ++q) fn(s, d.row_dofs[q], d.row_weights[q]);
      }
    }
  };
  d.t_offsets.assign(static_cast<std::size_t>(d.n_global) + 1, 0);
  visit([&](std::size_t, int32_t g, double) { ++d.t_offsets[static_cast<std::size_t>(g) + 1]; });
  for (int32_t g = 0; g < d.n_global; ++g) d.t_offsets[g + 1] += d.t_offsets[g];
  d.t_slots.resize(d.t_offsets.back());
  if (d.has_constraints()) d.t_weights.resize(d.t_offsets.back());
  std::vector<int32_t> fill(d.t_offsets.begin(), d.t_offsets.end() - 1);
  visit([&](std::size_t s, int32_t g, double w) {
    const int32_t at = fill[g]++;
    d.t_slots[at] = static_cast<int32_t>(s);
    if (!d.t_weights.empty()) d.t_weights[at] = w;
  });

  d.primary.assign(d.n_global, -1);
  for (std::size_t s = 0; s < nslots; ++s) {
    const int32_t m = d.l2g[s];
    if (m >= 0 && d.primary[m] < 0) d.primary[m] = static_cast<int32_t>(s);
  }
  for (int32_t g = 0; g < d.n_global; ++g)
    INSEM_REQUIRE(d.primary[g] >= 0, "global point " << g << " has no unconstrained element point");
  return d;
}

}  // namespace insem_amd
