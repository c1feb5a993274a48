// Local-to-global point map with hanging-node constraint rows and its
// transpose, built on the host.
//
// Restates ops::DofMap::build (cpp/src/ops/dof_map.cpp:42-260): global
// indices in first-touch order over the leaves (vertices, then edge interiors
// in canonical edge direction, then element interiors), constrained slots on
// hanging faces resolved recursively into rows of (global point, weight), and
// the transpose CSR with contributions in ascending slot order.  The
// ascending slot order is the assembly summation order of the hot path.
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
