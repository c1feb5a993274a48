#include "geometry.hpp"

#include <algorithm>
#include <cmath>

#include "error.hpp"

namespace insem_amd {

std::array<double, 2> leaf_map(const Forest& f, int32_t leaf, double xi, double eta) {
  const auto b = f.leaf_box(leaf);
  return f.macro().map(f.leaf(leaf).tree, b[0] + (xi + 1.0) * 0.5 * (b[1] - b[0]),
                       b[2] + (eta + 1.0) * 0.5 * (b[3] - b[2]));
}

std::array<double, 4> leaf_jacobian(const Forest& f, int32_t leaf, double xi, double eta) {
  const auto b = f.leaf_box(leaf);
  const double hx = 0.5 * (b[1] - b[0]), hy = 0.5 * (b[3] - b[2]);
  const auto j = f.macro().jacobian(f.leaf(leaf).tree, b[0] + (xi + 1.0) * hx, b[2] + (eta + 1.0) * hy);
  return {j[0] * hx, j[1] * hy, j[2] * hx, j[3] * hy};
}

Geometry compute_geometry(const Forest& f, const std::vector<double>& ref, const std::vector<double>& axial_xi) {
  INSEM_REQUIRE(ref.size() == axial_xi.size() && ref.size() >= 2, "reference point sets must have the same size >= 2");
  const MacroMesh& m = f.macro();
  Geometry g;
  g.ngll = static_cast<int>(ref.size());
  g.n_elements = f.n_leaves();
  const int ng = g.ngll;
  const std::size_t npe = static_cast<std::size_t>(ng) * ng;
  const std::size_t total = npe * g.n_elements;
  for (auto* v : {&g.x, &g.z, &g.xi_x, &g.xi_z, &g.eta_x, &g.eta_z, &g.jac}) v->resize(total);
  g.kind.resize(g.n_elements);
  constexpr double kTol = 1e-10;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < g.n_elements; ++e) {
    const int32_t cell = f.leaf(e).tree;
    const std::vector<double>& rx = f.is_axial(e) ? axial_xi : ref;
    const auto b = f.leaf_box(e);
    const double hx = 0.5 * (b[1] - b[0]), hy = 0.5 * (b[3] - b[2]);
    double scale = 0.0, dev = 0.0, off = 0.0;
    std::array<double, 4> first{};
    for (int jq = 0; jq < ng; ++jq) {
      for (int iq = 0; iq < ng; ++iq) {
        const double xi = b[0] + (rx[iq] + 1.0) * hx;
        const double eta = b[2] + (ref[jq] + 1.0) * hy;
        const auto x = m.map(cell, xi, eta);
        const auto jm = m.jacobian(cell, xi, eta);
        const std::array<double, 4> j{jm[0] * hx, jm[1] * hy, jm[2] * hx, jm[3] * hy};
        const double det = j[0] * j[3] - j[1] * j[2];
        const std::size_t p = e * npe + static_cast<std::size_t>(jq) * ng + iq;
        g.x[p] = x[0];
        g.z[p] = x[1];
        g.jac[p] = det;
        g.xi_x[p] = j[3] / det;
        g.xi_z[p] = -j[1] / det;
        g.eta_x[p] = -j[2] / det;
        g.eta_z[p] = j[0] / det;
        if (iq == 0 && jq == 0) first = j;
        for (int k = 0; k < 4; ++k) {
          scale = std::max(scale, std::fabs(j[k]));
          dev = std::max(dev, std::fabs(j[k] - first[k]));
        }
        off = std::max({off, std::fabs(j[1]), std::fabs(j[2])});
      }
    }
    g.kind[e] = dev > kTol * scale ? Kind::Curved : off > kTol * scale ? Kind::Affine : Kind::Rect;
  }
  for (double det : g.jac) INSEM_REQUIRE(det > 0.0, "leaf mapping has a non-positive Jacobian (" << det << ")");
  return g;
}

std::optional<Location> locate(const Forest& f, const Geometry& g, std::array<double, 2> x,
                               const std::vector<int32_t>& candidates) {
  const std::size_t npe = static_cast<std::size_t>(g.ngll) * g.ngll;
  for (int32_t leaf : candidates) {
    const std::size_t p0 = static_cast<std::size_t>(leaf) * npe;
    const auto [x0, x1] = std::minmax_element(g.x.begin() + p0, g.x.begin() + p0 + npe);
    const auto [z0, z1] = std::minmax_element(g.z.begin() + p0, g.z.begin() + p0 + npe);
    const double pad = 0.25 * std::max(*x1 - *x0, *z1 - *z0);
    if (x[0] < *x0 - pad || x[0] > *x1 + pad || x[1] < *z0 - pad || x[1] > *z1 + pad) continue;
    double xi = 0.0, eta = 0.0;
    bool ok = false;
    for (int it = 0; it < 30; ++it) {
      const auto y = leaf_map(f, leaf, xi, eta);
      const auto j = leaf_jacobian(f, leaf, xi, eta);
      const double rx = y[0] - x[0], rz = y[1] - x[1];
      const double det = j[0] * j[3] - j[1] * j[2];
      const double dxi = (j[3] * rx - j[1] * rz) / det;
      const double deta = (-j[2] * rx + j[0] * rz) / det;
      xi = std::clamp(xi - dxi, -2.0, 2.0);
      eta = std::clamp(eta - deta, -2.0, 2.0);
      if (std::fabs(dxi) + std::fabs(deta) < 1e-12) {
        ok = true;
        break;
      }
    }
    constexpr double kTol = 1e-9;
    if (ok && std::fabs(xi) <= 1.0 + kTol && std::fabs(eta) <= 1.0 + kTol)
      return Location{leaf, std::clamp(xi, -1.0, 1.0), std::clamp(eta, -1.0, 1.0)};
  }
  return std::nullopt;
}

PointWeights point_weights(const Forest& f, const Lagrange1D& bx, const Lagrange1D& by, const Location& loc) {
  const int n = bx.size();
  const auto lx = bx.values(loc.xi), ly = by.values(loc.eta);
  const auto dlx = bx.derivatives(loc.xi), dly = by.derivatives(loc.eta);
  const auto j = leaf_jacobian(f, loc.leaf, loc.xi, loc.eta);
  const double det = j[0] * j[3] - j[1] * j[2];
  const double xi_x = j[3] / det, xi_z = -j[1] / det, eta_x = -j[2] / det, eta_z = j[0] / det;
  PointWeights w;
  w.value.resize(static_cast<std::size_t>(n) * n);
  w.d_dx.resize(w.value.size());
  w.d_dz.resize(w.value.size());
  for (int jj = 0; jj < n; ++jj)
    for (int ii = 0; ii < n; ++ii) {
      const std::size_t k = static_cast<std::size_t>(jj) * n + ii;
      const double dxi = dlx[ii] * ly[jj], deta = lx[ii] * dly[jj];
      w.value[k] = lx[ii] * ly[jj];
      w.d_dx[k] = dxi * xi_x + deta * eta_x;
      w.d_dz[k] = dxi * xi_z + deta * eta_z;
    }
  return w;
}

}  // namespace insem_amd
