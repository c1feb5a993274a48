// Per-point geometry of the leaves and point location, on the host in
// double (the geometry of cpp/src/geometry/geometry.cpp:44-109 and the point
// location of cpp/src/solver/locate.cpp:9-75).
//
// Every leaf's map is its macro cell's map on the leaf's box; at each element
// point it gives the coordinates, the Jacobian determinant and the inverse
// metric.  A leaf is Rect when its Jacobian is one diagonal matrix, Affine
// when it is one matrix, Curved otherwise (relative tolerance 1e-10).  A
// point is located in the first candidate leaf whose map reaches it (Newton
// in reference coordinates) and weighted by the tensor Lagrange basis.
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <vector>

#include "basis.hpp"
#include "mesh.hpp"

namespace insem_amd {

enum class Kind : uint8_t { Rect = 0, Affine = 1, Curved = 2, Axisym = 3 };

struct Geometry {
  int ngll = 0;
  int32_t n_elements = 0;
  std::vector<double> x, z, xi_x, xi_z, eta_x, eta_z, jac;  // [leaf][k]
  std::vector<Kind> kind;                                    // per leaf
  // per leaf: the bounding box (min x, max x, min z, max z) of its points
  // (locate's cheap rejection; filled by compute_geometry)
  std::vector<std::array<double, 4>> box;
};

std::array<double, 2> leaf_map(const Forest& f, int32_t leaf, double xi, double eta);
std::array<double, 4> leaf_jacobian(const Forest& f, int32_t leaf, double xi, double eta);

// axial_xi: the xi points of axial leaves (GLJ) of an axisymmetric forest.
Geometry compute_geometry(const Forest& f, const std::vector<double>& ref, const std::vector<double>& axial_xi);

struct Location {
  int32_t leaf = -1;
  double xi = 0.0, eta = 0.0;
};

std::optional<Location> locate(const Forest& f, const Geometry& g, std::array<double, 2> x,
                               const std::vector<int32_t>& candidates);

struct PointWeights {
  std::vector<double> value, d_dx, d_dz;  // [j * ngll + i]
};

PointWeights point_weights(const Forest& f, const Lagrange1D& bx, const Lagrange1D& by, const Location& loc);

}  // namespace insem_amd
