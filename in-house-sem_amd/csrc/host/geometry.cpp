#include "geometry.hpp"

#include <algorithm>
#include <cmath>

#include "error.hpp"

namespace insem_amd {
namespace {

// A leaf's map from its reference square [-1, 1]^2 into the plane: the
// macro cell's map restricted to the leaf's box of the cell's reference
// square.
class LeafMap {
 public:
  LeafMap(const Forest& f, int32_t leaf) : macro_(f.macro()), cell_(f.leaf(leaf).tree) {
    const auto b = f.leaf_box(leaf);
    x0_ = b[0];
    y0_ = b[2];
    hx_ = 0.5 * (b[1] - b[0]);
    hy_ = 0.5 * (b[3] - b[2]);
  }
  double cell_xi(double xi) const { return x0_ + (xi + 1.0) * hx_; }
  double cell_eta(double eta) const { return y0_ + (eta + 1.0) * hy_; }
  std::array<double, 2> point(double xi, double eta) const { return macro_.map(cell_, cell_xi(xi), cell_eta(eta)); }
  // {dx/dxi, dx/deta, dz/dxi, dz/deta} of the leaf's own coordinates
  std::array<double, 4> jacobian(double xi, double eta) const {
    const auto c = macro_.jacobian(cell_, cell_xi(xi), cell_eta(eta));
    return {c[0] * hx_, c[1] * hy_, c[2] * hx_, c[3] * hy_};
  }

 private:
  const MacroMesh& macro_;
  int32_t cell_;
  double x0_, y0_, hx_, hy_;
};

// The inverse metric of a Jacobian: (xi_x, xi_z, eta_x, eta_z) and det.
struct InverseMetric {
  double xi_x, xi_z, eta_x, eta_z, det;
};

InverseMetric invert(const std::array<double, 4>& j) {
  const double det = j[0] * j[3] - j[1] * j[2];
  return {j[3] / det, -j[1] / det, -j[2] / det, j[0] / det, det};
}

// Rect: the Jacobian is the same diagonal matrix at every point; Affine:
// the same matrix; Curved: otherwise.  Deviations are measured relative to
// the largest Jacobian entry of the leaf.
class KindVote {
 public:
  void add(const std::array<double, 4>& j) {
    if (!seen_) {
      ref_ = j;
      seen_ = true;
    }
    for (int k = 0; k < 4; ++k) {
      largest_ = std::max(largest_, std::fabs(j[k]));
      spread_ = std::max(spread_, std::fabs(j[k] - ref_[k]));
    }
    shear_ = std::max(shear_, std::max(std::fabs(j[1]), std::fabs(j[2])));
  }
  Kind kind() const {
    constexpr double kRel = 1e-10;
    if (spread_ > kRel * largest_) return Kind::Curved;
    return shear_ > kRel * largest_ ? Kind::Affine : Kind::Rect;
  }

 private:
  bool seen_ = false;
  std::array<double, 4> ref_{};
  double largest_ = 0.0, spread_ = 0.0, shear_ = 0.0;
};

// The bounding box of leaf e's points.
std::array<double, 4> point_box(const Geometry& g, int32_t e) {
  const std::size_t npe = static_cast<std::size_t>(g.ngll) * g.ngll;
  const std::size_t p0 = static_cast<std::size_t>(e) * npe;
  double lo_x = g.x[p0], hi_x = lo_x, lo_z = g.z[p0], hi_z = lo_z;
  for (std::size_t p = p0; p < p0 + npe; ++p) {
    lo_x = std::min(lo_x, g.x[p]), hi_x = std::max(hi_x, g.x[p]);
    lo_z = std::min(lo_z, g.z[p]), hi_z = std::max(hi_z, g.z[p]);
  }
  return {lo_x, hi_x, lo_z, hi_z};
}

}  // namespace

std::array<double, 2> leaf_map(const Forest& f, int32_t leaf, double xi, double eta) {
  return LeafMap(f, leaf).point(xi, eta);
}

std::array<double, 4> leaf_jacobian(const Forest& f, int32_t leaf, double xi, double eta) {
  return LeafMap(f, leaf).jacobian(xi, eta);
}

Geometry compute_geometry(const Forest& f, const std::vector<double>& ref, const std::vector<double>& axial_xi) {
  INSEM_REQUIRE(ref.size() >= 2 && axial_xi.size() == ref.size(),
                "geometry needs two point sets of one size (at least 2), got " << ref.size() << " and "
                                                                                << axial_xi.size());
  Geometry g;
  g.ngll = static_cast<int>(ref.size());
  g.n_elements = f.n_leaves();
  const int ng = g.ngll;
  const std::size_t npe = static_cast<std::size_t>(ng) * ng;
  for (auto* v : {&g.x, &g.z, &g.xi_x, &g.xi_z, &g.eta_x, &g.eta_z, &g.jac}) v->assign(npe * g.n_elements, 0.0);
  g.kind.assign(static_cast<std::size_t>(g.n_elements), Kind::Rect);
  g.box.assign(static_cast<std::size_t>(g.n_elements), {0.0, 0.0, 0.0, 0.0});
  int32_t bad_leaf = -1;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < g.n_elements; ++e) {
    const LeafMap map(f, e);
    // leaves on the axis carry the GLJ points along xi
    const std::vector<double>& xs = f.is_axial(e) ? axial_xi : ref;
    KindVote vote;
    std::size_t p = static_cast<std::size_t>(e) * npe;
    for (int jq = 0; jq < ng; ++jq)
      for (int iq = 0; iq < ng; ++iq, ++p) {
        const auto pos = map.point(xs[iq], ref[jq]);
        const auto jm = map.jacobian(xs[iq], ref[jq]);
        const InverseMetric im = invert(jm);
        g.x[p] = pos[0];
        g.z[p] = pos[1];
        g.jac[p] = im.det;
        g.xi_x[p] = im.xi_x;
        g.xi_z[p] = im.xi_z;
        g.eta_x[p] = im.eta_x;
        g.eta_z[p] = im.eta_z;
        vote.add(jm);
        if (!(im.det > 0.0)) {
#pragma omp critical
          bad_leaf = e;
        }
      }
    g.kind[e] = vote.kind();
    g.box[e] = point_box(g, e);
  }
  INSEM_REQUIRE(bad_leaf < 0, "leaf " << bad_leaf << " maps onto the plane with a non-positive Jacobian");
  return g;
}

std::optional<Location> locate(const Forest& f, const Geometry& g, std::array<double, 2> x,
                               const std::vector<int32_t>& candidates) {
  const bool boxed = g.box.size() == static_cast<std::size_t>(g.n_elements);
  for (int32_t leaf : candidates) {
    // cheap rejection: the leaf's points' bounding box, widened by a quarter
    const std::array<double, 4> b = boxed ? g.box[leaf] : point_box(g, leaf);
    const double lo_x = b[0], hi_x = b[1], lo_z = b[2], hi_z = b[3];
    const double margin = 0.25 * std::max(hi_x - lo_x, hi_z - lo_z);
    if (x[0] < lo_x - margin || x[0] > hi_x + margin || x[1] < lo_z - margin || x[1] > hi_z + margin) continue;
    // Newton on map(xi, eta) = x from the leaf's centre, iterates kept in
    // [-2, 2]^2
    const LeafMap map(f, leaf);
    double xi = 0.0, eta = 0.0;
    bool converged = false;
    for (int it = 0; it < 30 && !converged; ++it) {
      const auto y = map.point(xi, eta);
      const InverseMetric im = invert(map.jacobian(xi, eta));
      const double rx = y[0] - x[0], rz = y[1] - x[1];
      // step = J^-1 (map - x)
      const double sxi = im.xi_x * rx + im.xi_z * rz, seta = im.eta_x * rx + im.eta_z * rz;
      xi = std::clamp(xi - sxi, -2.0, 2.0);
      eta = std::clamp(eta - seta, -2.0, 2.0);
      converged = std::fabs(sxi) + std::fabs(seta) < 1e-12;
    }
    // the first candidate that holds the point (to 1e-9 in reference
    // coordinates, so points on shared edges go to the earlier leaf)
    constexpr double kInside = 1.0 + 1e-9;
    if (converged && std::fabs(xi) <= kInside && std::fabs(eta) <= kInside)
      return Location{leaf, std::clamp(xi, -1.0, 1.0), std::clamp(eta, -1.0, 1.0)};
  }
  return std::nullopt;
}

PointWeights point_weights(const Forest& f, const Lagrange1D& bx, const Lagrange1D& by, const Location& loc) {
  const int n = bx.size();
  const std::vector<double> vx = bx.values(loc.xi), vy = by.values(loc.eta);
  const std::vector<double> dx = bx.derivatives(loc.xi), dy = by.derivatives(loc.eta);
  const InverseMetric im = invert(leaf_jacobian(f, loc.leaf, loc.xi, loc.eta));
  PointWeights w;
  const std::size_t npe = static_cast<std::size_t>(n) * n;
  w.value.resize(npe);
  w.d_dx.resize(npe);
  w.d_dz.resize(npe);
  std::size_t k = 0;
  for (int jj = 0; jj < n; ++jj)
    for (int ii = 0; ii < n; ++ii, ++k) {
      // reference derivatives, then the chain rule through the inverse metric
      const double along_xi = dx[ii] * vy[jj], along_eta = vx[ii] * dy[jj];
      w.value[k] = vx[ii] * vy[jj];
      w.d_dx[k] = along_xi * im.xi_x + along_eta * im.eta_x;
      w.d_dz[k] = along_xi * im.xi_z + along_eta * im.eta_z;
    }
  return w;
}

}  // namespace insem_amd
