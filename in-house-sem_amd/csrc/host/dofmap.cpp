#include "dofmap.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <unordered_map>

#include "basis.hpp"
#include "error.hpp"

namespace insem_amd {
namespace {

constexpr int32_t kPending = std::numeric_limits<int32_t>::min();  // on a hanging face, not yet resolved

// Slot (i, j) of an element with n intervals per direction lies on corner
// 0..3, on the interior of face 0..3 at position `along` in the face's
// counter-clockwise direction, or inside the element.
struct SlotPlace {
  int corner = -1;
  int face = -1;
  int along = 0;
};

SlotPlace place_of(int i, int j, int n) {
  SlotPlace p;
  const bool xe = i == 0 || i == n, ye = j == 0 || j == n;
  if (xe && ye) {
    p.corner = j == 0 ? (i == 0 ? 0 : 1) : (i == n ? 2 : 3);
  } else if (j == 0) {
    p.face = 0, p.along = i;
  } else if (i == n) {
    p.face = 1, p.along = j;
  } else if (j == n) {
    p.face = 2, p.along = n - i;
  } else if (i == 0) {
    p.face = 3, p.along = n - j;
  }
  return p;
}

bool on_face(int f, int i, int j, int n) {
  return (f == 0 && j == 0) || (f == 1 && i == n) || (f == 2 && j == n) || (f == 3 && i == 0);
}

// Slot index of the point m along face f, m counted in the element's own i
// (faces 0, 2) or j (faces 1, 3) direction.
int face_slot(int f, int m, int ngll) {
  const int n = ngll - 1;
  switch (f) {
    case 0: return m;
    case 1: return m * ngll + n;
    case 2: return n * ngll + m;
    default: return m * ngll;
  }
}

using Row = std::vector<std::pair<int32_t, double>>;

// Adds w * entries into a row kept sorted by global index.
void accumulate(Row& row, const Row& entries, double w) {
  for (const auto& [g, wg] : entries) {
    auto it = std::lower_bound(row.begin(), row.end(), g,
                               [](const std::pair<int32_t, double>& a, int32_t b) { return a.first < b; });
    if (it != row.end() && it->first == g)
      it->second += w * wg;
    else
      row.insert(it, {g, w * wg});
  }
}

class Numbering {
 public:
  Numbering(const Forest& forest, int ngll, const std::vector<int32_t>& leaves)
      : forest_(forest), leaves_(leaves), gll_(gll_rule(ngll).points), glj_(glj01_rule(ngll).points) {
    d_.ngll = ngll;
    d_.npe = ngll * ngll;
    d_.n_elements = static_cast<int32_t>(leaves.size());
    n_ = ngll - 1;
    nslots_ = static_cast<std::size_t>(d_.n_elements) * d_.npe;
    INSEM_REQUIRE(nslots_ < static_cast<std::size_t>(std::numeric_limits<int32_t>::max()),
                  nslots_ << " element points do not fit 32-bit slot indices");
    element_of_leaf_.assign(static_cast<std::size_t>(forest.n_leaves()), -1);
    for (int32_t e = 0; e < d_.n_elements; ++e) {
      INSEM_REQUIRE(leaves[e] >= 0 && leaves[e] < forest.n_leaves(), "element " << e << " names leaf " << leaves[e]
                                                                               << ", which the forest does not have");
      element_of_leaf_[leaves[e]] = e;
    }
    corners_ = forest.corners();
  }

  DofMap run() {
    find_hanging_faces();
    number_free_points();
    build_rows();
    build_transpose();
    return std::move(d_);
  }

 private:
  struct Hanging {
    int32_t coarse = -1;  // element whose face this face lies on
    int coarse_face = -1;
  };
  struct Edge {
    int32_t high;   // the edge's other vertex (the higher id)
    int32_t first;  // global index of its first interior point
    int uses;
  };

  const int32_t* corners_of(int32_t e) const { return &corners_.corner[4 * static_cast<std::size_t>(leaves_[e])]; }

  // Faces of this domain's elements that lie on half of a coarser
  // element's face, and the vertices at the middle of such coarse faces.
  void find_hanging_faces() {
    hanging_.assign(static_cast<std::size_t>(d_.n_elements), {});
    middle_.assign(static_cast<std::size_t>(corners_.n_vertices), 0);
    for (int32_t e = 0; e < d_.n_elements; ++e) {
      for (int f = 0; f < 4; ++f) {
        const Neighbor nb = forest_.neighbor(leaves_[e], f);
        if (nb.leaf < 0) continue;
        const int32_t other = element_of_leaf_[nb.leaf];
        if (other < 0) {
          INSEM_REQUIRE(nb.level_diff == 0, "a face between two media joins leaves of refinement levels "
                                                << forest_.leaf(leaves_[e]).level << " and "
                                                << forest_.leaf(nb.leaf).level << "; media must meet conformingly");
          continue;
        }
        if (nb.level_diff >= 0) continue;
        INSEM_REQUIRE(nb.level_diff == -1, "leaf " << leaves_[e] << " meets a leaf " << -nb.level_diff
                                                   << " levels coarser: the forest must be 2:1 balanced");
        hanging_[e][f] = {other, nb.face};
        const int32_t* cc = corners_of(other);
        const int32_t a = cc[kFaceCorner[nb.face][0]], b = cc[kFaceCorner[nb.face][1]];
        for (int s = 0; s < 2; ++s) {
          const int32_t v = corners_of(e)[kFaceCorner[f][s]];
          if (v != a && v != b) middle_[v] = 1;
        }
      }
    }
  }

  // The edge between vertices a and b: its first interior index, reserving
  // n - 1 new indices the first time it is met.
  int32_t edge_first(int32_t a, int32_t b) {
    const int32_t lo = std::min(a, b), hi = std::max(a, b);
    std::vector<Edge>& list = edges_[lo];
    for (Edge& ed : list)
      if (ed.high == hi) {
        INSEM_REQUIRE(++ed.uses <= 2, "edge (" << lo << ", " << hi << ") belongs to more than two leaves");
        return ed.first;
      }
    list.push_back({hi, next_, 1});
    next_ += n_ - 1;
    return list.back().first;
  }

  // First-touch numbering over the elements in order: an element's new
  // edges take their interior indices first (faces in order, each edge
  // numbered from its lower vertex to its higher), then its slots, j outer
  // and i inner, give new vertices and interior points theirs.  Slots on
  // hanging faces and middle vertices wait for their constraint rows.
  void number_free_points() {
    d_.l2g.assign(nslots_, -1);
    std::vector<int32_t> vertex_index(static_cast<std::size_t>(corners_.n_vertices), -1);
    edges_.clear();
    next_ = 0;
    for (int32_t e = 0; e < d_.n_elements; ++e) {
      const int32_t* v = corners_of(e);
      int32_t first[4];
      bool low_first[4];
      for (int f = 0; f < 4; ++f) {
        const int32_t a = v[kFaceCorner[f][0]], b = v[kFaceCorner[f][1]];
        low_first[f] = a < b;
        first[f] = hanging_[e][f].coarse >= 0 ? -1 : edge_first(a, b);
      }
      int32_t* slot = &d_.l2g[static_cast<std::size_t>(e) * d_.npe];
      for (int j = 0; j <= n_; ++j)
        for (int i = 0; i <= n_; ++i) {
          const SlotPlace p = place_of(i, j, n_);
          int32_t g;
          if (p.corner >= 0) {
            const int32_t vert = v[p.corner];
            if (middle_[vert]) {
              g = kPending;
            } else {
              if (vertex_index[vert] < 0) vertex_index[vert] = next_++;
              g = vertex_index[vert];
            }
          } else if (p.face >= 0) {
            g = first[p.face] < 0 ? kPending : first[p.face] + (low_first[p.face] ? p.along : n_ - p.along) - 1;
          } else {
            g = next_++;
          }
          slot[j * d_.ngll + i] = g;
        }
    }
    d_.n_global = next_;
  }

  const std::vector<double>& face_nodes(int32_t e, int f) const {
    return forest_.is_axial(leaves_[e]) && (f == 0 || f == 2) ? glj_.nodes() : gll_.nodes();
  }
  const Lagrange1D& face_basis(int32_t e, int f) const {
    return forest_.is_axial(leaves_[e]) && (f == 0 || f == 2) ? glj_ : gll_;
  }

  // Global entries of a slot: itself, its finished row, or (for a slot on
  // a hanging face) the coarse face's interpolant at the slot, each coarse
  // slot replaced by its own entries (coarse slots may hang in turn).
  Row entries_of(std::size_t s) {
    const int32_t m = d_.l2g[s];
    if (m >= 0) return {{m, 1.0}};
    if (m != kPending) {
      const int32_t r = -1 - m;
      Row out;
      for (int32_t q = d_.row_offsets[r]; q < d_.row_offsets[r + 1]; ++q) out.push_back({d_.row_dofs[q], d_.row_weights[q]});
      return out;
    }
    const auto done = pending_rows_.find(s);
    if (done != pending_rows_.end()) return done->second;
    const int32_t e = static_cast<int32_t>(s / d_.npe);
    const int k = static_cast<int>(s % d_.npe);
    const int i = k % d_.ngll, j = k / d_.ngll;
    int f = 0;
    while (f < 4 && !(hanging_[e][f].coarse >= 0 && on_face(f, i, j, n_))) ++f;
    INSEM_REQUIRE(f < 4, "slot " << k << " of element " << e << " waits for a constraint but lies on no hanging face");
    const Hanging h = hanging_[e][f];
    const double t = face_nodes(e, f)[(f == 0 || f == 2) ? i : j];
    const double tc = forest_.face_coordinate(leaves_[e], f, t, leaves_[h.coarse], h.coarse_face);
    const std::vector<double> w = face_basis(h.coarse, h.coarse_face).values(tc);
    Row row;
    for (int q = 0; q <= n_; ++q) {
      if (std::fabs(w[q]) < 1e-14) continue;
      const std::size_t cs = static_cast<std::size_t>(h.coarse) * d_.npe + face_slot(h.coarse_face, q, d_.ngll);
      accumulate(row, entries_of(cs), w[q]);
    }
    pending_rows_.emplace(s, row);
    return row;
  }

  // Each waiting slot, in slot order, becomes a direct index (a single
  // entry of weight 1) or the next constraint row.
  void build_rows() {
    d_.row_offsets.assign(1, 0);
    for (std::size_t s = 0; s < nslots_; ++s) {
      if (d_.l2g[s] != kPending) continue;
      const Row row = entries_of(s);
      if (row.size() == 1 && std::fabs(row[0].second - 1.0) < 1e-12) {
        d_.l2g[s] = row[0].first;
        continue;
      }
      d_.l2g[s] = -1 - d_.n_rows();
      for (const auto& [g, w] : row) {
        d_.row_dofs.push_back(g);
        d_.row_weights.push_back(w);
      }
      d_.row_offsets.push_back(static_cast<int32_t>(d_.row_dofs.size()));
    }
  }

  // For every global point its contributing slots in ascending slot order
  // (with the constraint weight when the map has rows), and the first
  // unconstrained slot of each point.
  void build_transpose() {
    // every (point, slot, weight) contribution, visited in ascending slot order
    auto each_contribution = [&](auto&& use) {
      for (std::size_t s = 0; s < nslots_; ++s) {
        const int32_t m = d_.l2g[s];
        if (m >= 0) {
          use(m, static_cast<int32_t>(s), 1.0);
          continue;
        }
        for (int32_t q = d_.row_offsets[-1 - m]; q < d_.row_offsets[-m]; ++q)
          use(d_.row_dofs[q], static_cast<int32_t>(s), d_.row_weights[q]);
      }
    };
    std::vector<int32_t> count(static_cast<std::size_t>(d_.n_global) + 1, 0);
    each_contribution([&](int32_t g, int32_t, double) { ++count[static_cast<std::size_t>(g) + 1]; });
    for (int32_t g = 0; g < d_.n_global; ++g) count[g + 1] += count[g];
    d_.t_offsets = count;
    d_.t_slots.resize(static_cast<std::size_t>(count.back()));
    if (d_.has_constraints()) d_.t_weights.resize(d_.t_slots.size());
    each_contribution([&](int32_t g, int32_t slot, double w) {
      const int32_t at = count[g]++;
      d_.t_slots[at] = slot;
      if (!d_.t_weights.empty()) d_.t_weights[at] = w;
    });

    d_.primary.assign(static_cast<std::size_t>(d_.n_global), -1);
    for (std::size_t s = nslots_; s-- > 0;) {
      const int32_t m = d_.l2g[s];
      if (m >= 0) d_.primary[m] = static_cast<int32_t>(s);
    }
    for (int32_t g = 0; g < d_.n_global; ++g)
      INSEM_REQUIRE(d_.primary[g] >= 0, "global point " << g << " is reached only through constraint rows");
  }

  const Forest& forest_;
  const std::vector<int32_t>& leaves_;
  const Lagrange1D gll_, glj_;
  DofMap d_;
  int n_ = 0;
  std::size_t nslots_ = 0;
  std::vector<int32_t> element_of_leaf_;
  Corners corners_;
  std::vector<std::array<Hanging, 4>> hanging_;
  std::vector<char> middle_;
  std::unordered_map<int32_t, std::vector<Edge>> edges_;
  int32_t next_ = 0;
  std::unordered_map<std::size_t, Row> pending_rows_;
};

}  // namespace

int32_t DofMap::global_index(int32_t e, int k) const {
  const int32_t m = l2g[static_cast<std::size_t>(e) * npe + k];
  INSEM_REQUIRE(m >= 0, "point " << k << " of element " << e << " is constrained (it lies on a hanging face)");
  return m;
}

std::vector<std::pair<int32_t, double>> DofMap::slot_entries(int32_t e, int k) const {
  const int32_t m = l2g[static_cast<std::size_t>(e) * npe + k];
  if (m >= 0) return {{m, 1.0}};
  std::vector<std::pair<int32_t, double>> out;
  for (int32_t q = row_offsets[-1 - m]; q < row_offsets[-m]; ++q) out.push_back({row_dofs[q], row_weights[q]});
  return out;
}

std::vector<int32_t> DofMap::multiplicity() const {
  std::vector<int32_t> m(static_cast<std::size_t>(n_global));
  for (int32_t g = 0; g < n_global; ++g) m[g] = t_offsets[g + 1] - t_offsets[g];
  return m;
}

DofMap DofMap::build(const Forest& forest, int ngll, const std::vector<int32_t>& leaves) {
  INSEM_REQUIRE(ngll >= 2, "an element needs at least 2 points per direction, not " << ngll);
  return Numbering(forest, ngll, leaves).run();
}

}  // namespace insem_amd
