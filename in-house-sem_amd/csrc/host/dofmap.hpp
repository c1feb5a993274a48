// Local-to-global point map of one medium, its hanging-node constraint rows
// and its transpose, built on the host (the map of
// cpp/include/insem/ops/dof_map.hpp:30-92, with the same numbering).
//
// Numbering: elements in leaf order; an element's new edges take their
// interior indices first (faces in order, each edge counted from its lower
// vertex id), then its slots (j outer, i inner) give new vertices and
// interior points theirs.  Slots on faces that lie on half of a coarser face
// are the coarse face's interpolant at the slot: a direct index when that is
// one point with weight 1, otherwise a constraint row of (point, weight)
// pairs in ascending point order.  The transpose lists every point's
// contributing slots in ascending slot order, the assembly summation order
// of the hot path.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "mesh.hpp"

namespace insem_amd {

struct DofMap {
  int ngll = 0;
  int npe = 0;
  int32_t n_elements = 0;
  int32_t n_global = 0;
  std::vector<int32_t> l2g;               // [e * npe + k]: >= 0 global point, < 0 row -1 - l2g
  std::vector<int32_t> row_offsets{0};
  std::vector<int32_t> row_dofs;
  std::vector<double> row_weights;
  std::vector<int32_t> primary;           // an unconstrained slot of each global point
  std::vector<int32_t> t_offsets;
  std::vector<int32_t> t_slots;
  std::vector<double> t_weights;          // empty without constraints

  int32_t n_rows() const { return static_cast<int32_t>(row_offsets.size()) - 1; }
  bool has_constraints() const { return n_rows() > 0; }
  int32_t global_index(int32_t e, int k) const;
  std::vector<std::pair<int32_t, double>> slot_entries(int32_t e, int k) const;
  std::vector<int32_t> multiplicity() const;

  static DofMap build(const Forest& forest, int ngll, const std::vector<int32_t>& leaves);
};

}  // namespace insem_amd
