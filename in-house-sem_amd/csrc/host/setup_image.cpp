// The time-loop state of one model as plain arrays: what the reference's
// Solver constructor builds (cpp/src/solver/solver.cpp:147-568), produced by
// ImageBuilder in stages -- kernel tables, one medium's elements, blocks and
// mass, boundary terms, the fluid-solid coupling, sources, receivers.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>

#include "error.hpp"
#include "model.hpp"

namespace insem_amd {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtHalf = 0.70710678118654752440;

// (i, j) of the point q of face f, counted counter-clockwise round the element.
std::pair<int, int> face_point(int f, int q, int n) {
  switch (f) {
    case 0: return {q, 0};
    case 1: return {n, q};
    case 2: return {n - q, n};
    default: return {0, n - q};
  }
}

// Per-point sums over the points of one medium, in double, keeping the
// points touched in the order first touched; `ascending()` lists them by
// point index (the order the reference's ordered maps give).
template <int W>
class PointSums {
 public:
  explicit PointSums(int32_t n_points) : sum_(static_cast<std::size_t>(n_points) * W, 0.0), hit_(n_points, 0) {}
  double* at(int32_t g) {
    if (!hit_[g]) {
      hit_[g] = 1;
      touched_.push_back(g);
    }
    return &sum_[static_cast<std::size_t>(g) * W];
  }
  const double* get(int32_t g) const { return &sum_[static_cast<std::size_t>(g) * W]; }
  std::vector<int32_t> ascending() const {
    std::vector<int32_t> out = touched_;
    std::sort(out.begin(), out.end());
    return out;
  }

 private:
  std::vector<double> sum_;
  std::vector<char> hit_;
  std::vector<int32_t> touched_;
};

// The same sums for the few points one source or receiver touches (tens,
// where the medium has millions): a short list searched from its end, so
// that no medium-sized array is cleared per source or receiver.
template <int W>
class FewPointSums {
 public:
  double* at(int32_t g) {
    for (std::size_t k = points_.size(); k-- > 0;)
      if (points_[k] == g) return sums_[k].data();
    points_.push_back(g);
    sums_.push_back({});
    return sums_.back().data();
  }
  const double* get(int32_t g) const {
    std::size_t k = 0;
    while (k < points_.size() && points_[k] != g) ++k;
    INSEM_REQUIRE(k < points_.size(), "point " << g << " has no sum");
    return sums_[k].data();
  }
  std::vector<int32_t> ascending() const {
    std::vector<int32_t> out = points_;
    std::sort(out.begin(), out.end());
    return out;
  }

 private:
  std::vector<int32_t> points_;
  std::vector<std::array<double, W>> sums_;
};

class ImageBuilder {
 public:
  ImageBuilder(const Forest& F, const Geometry& G, const Rule& gll, int azimuthal_order,
               const std::map<int, Material>& materials, const std::map<int, Boundary>& boundaries)
      : F_(F),
        G_(G),
        M_(F.macro()),
        gll_(gll),
        glj_(glj01_rule(gll.size())),
        lag_(gll.points),
        lag_glj_(glj_.points),
        materials_(materials),
        boundaries_(boundaries),
        ngll_(gll.size()),
        n_(gll.size() - 1),
        npe_(static_cast<std::size_t>(gll.size()) * gll.size()),
        axisym_(F.macro().axisymmetric()),
        am_(azimuthal_order) {}

  std::unique_ptr<SetupImage> build(const TimeConfig& time, const std::vector<Source>& sources,
                                    const std::vector<Receiver>& receivers) {
    INSEM_REQUIRE(am_ == 0 || (am_ == 1 && axisym_), "azimuthal order " << am_ << " runs on axisymmetric meshes only");
    INSEM_REQUIRE(ngll_ >= 2 && ngll_ <= 9, "the kernels take polynomial orders 1 to 8, not " << ngll_ - 1);
    INSEM_REQUIRE(time.record_every >= 1 && time.energy_every >= 0,
                  "record_every must be at least 1 and energy_every at least 0");
    INSEM_REQUIRE(time.snapshot_every >= 0, "snapshot_every must be at least 0, not " << time.snapshot_every);
    if (time.snapshot_every > 0)
      for (const std::string& q : time.snapshot_quantities)
        INSEM_REQUIRE(q == "displacement" || q == "velocity" || q == "acceleration" || q == "pressure",
                      "snapshots record displacement, velocity, acceleration or pressure, not '" << q << "'");
    snapshots_ = time.snapshot_every > 0;
    for (int32_t c = 0; c < M_.n_cells(); ++c) {
      const auto it = materials_.find(M_.material(c));
      INSEM_REQUIRE(it != materials_.end(), "region " << M_.material(c) << " has no material");
      it->second.validate();
    }
    img_ = std::make_unique<SetupImage>();
    img_->ngll = ngll_;
    img_->axisym = axisym_;
    img_->azimuthal_order = am_;
    img_->time = time;
    kernel_tables();
    for (Medium m : {Medium::Acoustic, Medium::Elastic}) add_domain(m);
    for (auto& d : img_->domains) boundary_terms(d);
    coupling();
    for (std::size_t s = 0; s < sources.size(); ++s) add_source(s, sources[s]);
    for (const Receiver& r : receivers) add_receiver(r);
    img_->refresh_views();
    return std::move(img_);
  }

 private:
  bool snapshots_ = false;  // keep the element points' coordinates
  const Material& material(int32_t leaf) const { return materials_.at(M_.material(F_.leaf(leaf).tree)); }
  std::size_t point_of(int32_t leaf, int i, int j) const {
    return static_cast<std::size_t>(leaf) * npe_ + static_cast<std::size_t>(j) * ngll_ + i;
  }

  // Derivative matrices on GLL and GLJ points and the tensor weights
  // w2[j * ngll + i] = w_i w_j (kernels.cpp:472-485).
  void kernel_tables() {
    auto as_float = [](const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); };
    img_->D = as_float(lag_.derivative_matrix());
    img_->Dj = as_float(lag_glj_.derivative_matrix());
    img_->w2.resize(npe_);
    for (int j = 0; j < ngll_; ++j)
      for (int i = 0; i < ngll_; ++i) img_->w2[j * ngll_ + i] = static_cast<float>(gll_.weights[i] * gll_.weights[j]);
  }

  // Quadrature weight of element point (i, j) with the Jacobian, and (on
  // axisymmetric meshes) with the radius: wr for the stiffness and mass,
  // hoop for the 1/r terms.  Along xi of an axial leaf the GLJ rule of the
  // weight (1 + xi) carries r: r / (1 + xi_i), or dr/dxi at the axis node.
  void point_weights_of(int32_t leaf, bool axial, int i, int j, double& wr, double& hoop) const {
    const std::size_t p = point_of(leaf, i, j);
    const double jac = G_.jac[p];
    if (!axisym_) {
      wr = gll_.weights[i] * gll_.weights[j] * jac;
      hoop = 0.0;
    } else if (!axial) {
      hoop = gll_.weights[i] * gll_.weights[j] * jac;
      wr = hoop * std::max(G_.x[p], 0.0);
    } else {
      const double r_over = i == 0 ? G_.eta_z[p] * jac : G_.x[p] / (1.0 + glj_.points[i]);
      wr = glj_.weights[i] * gll_.weights[j] * jac * r_over;
      hoop = i == 0 ? 0.0 : wr / G_.x[p];
    }
  }

  // One medium: its leaves in forest order, point map, element blocks,
  // lumped mass, coordinates and axis points (solver.cpp:147-254).
  void add_domain(Medium medium) {
    SetupImage::Domain d;
    d.medium = medium;
    d.azimuthal_order = am_;
    const bool elastic = medium == Medium::Elastic;
    d.ncomp = elastic ? (am_ == 1 ? 3 : 2) : 1;
    // components that vanish on the axis: u_r (m = 0); U+ and U_z (m = 1); P (m = 1)
    if (axisym_) {
      if (elastic)
        d.axis_components = am_ == 1 ? std::vector<int32_t>{0, 2} : std::vector<int32_t>{0};
      else if (am_ == 1)
        d.axis_components = {0};
    }
    d.element_of_leaf.assign(static_cast<std::size_t>(F_.n_leaves()), -1);
    for (int32_t l = 0; l < F_.n_leaves(); ++l)
      if (material(l).medium == medium) {
        d.element_of_leaf[l] = static_cast<int32_t>(d.leaves.size());
        d.leaves.push_back(l);
      }
    if (d.leaves.empty()) return;
    d.dofs = DofMap::build(F_, ngll_, d.leaves);

    const std::size_t ne = d.leaves.size();
    std::vector<double> wr(ne * npe_), hoop(ne * npe_), coef0(ne * npe_), coef1(ne * npe_), slot_mass(ne * npe_);
    std::vector<uint8_t> axial(ne, 0);
    for (std::size_t e = 0; e < ne; ++e) {
      const int32_t leaf = d.leaves[e];
      axial[e] = axisym_ && F_.is_axial(leaf) ? 1 : 0;
      const Material& m = material(leaf);
      // elastic: lambda, mu and rho; acoustic: 1/rho and 1/kappa
      const double c0 = elastic ? m.rho * (m.vp * m.vp - 2.0 * m.vs * m.vs) : 1.0 / m.rho;
      const double c1 = elastic ? m.rho * m.vs * m.vs : 0.0;
      for (int j = 0; j < ngll_; ++j)
        for (int i = 0; i < ngll_; ++i) {
          const std::size_t q = e * npe_ + static_cast<std::size_t>(j) * ngll_ + i;
          point_weights_of(leaf, axial[e] != 0, i, j, wr[q], hoop[q]);
          coef0[q] = c0;
          coef1[q] = c1;
          slot_mass[q] = elastic ? wr[q] * m.rho : wr[q] / (m.rho * m.vp * m.vp);
        }
    }
    if (axisym_)
      d.blocks.push_back(axisym_block(d, axial, coef0, coef1, wr, hoop));
    else
      for (Kind k : {Kind::Rect, Kind::Affine, Kind::Curved}) {
        SetupImage::Block b = cartesian_block(d, k, coef0, coef1);
        if (!b.elements.empty()) d.blocks.push_back(std::move(b));
      }
    lumped_mass(d, slot_mass);
    const int32_t ng = d.dofs.n_global;
    d.coords.resize(2 * static_cast<std::size_t>(ng));
    for (int32_t g = 0; g < ng; ++g) {
      const std::size_t s = static_cast<std::size_t>(d.dofs.primary[g]);
      const std::size_t p = static_cast<std::size_t>(d.leaves[s / npe_]) * npe_ + s % npe_;
      d.coords[g] = G_.x[p];
      d.coords[ng + g] = G_.z[p];
      if (!d.axis_components.empty() && G_.x[p] <= 0.0) d.axis_points.push_back(g);
    }
    if (snapshots_) {
      d.point_x.resize(ne * npe_);
      d.point_z.resize(ne * npe_);
      for (std::size_t e = 0; e < ne; ++e)
        for (std::size_t k = 0; k < npe_; ++k) {
          const std::size_t p = static_cast<std::size_t>(d.leaves[e]) * npe_ + k;
          d.point_x[e * npe_ + k] = G_.x[p];
          d.point_z[e * npe_ + k] = G_.z[p];
        }
    }
    img_->domains.push_back(std::move(d));
  }

  // One Axisym block of all elements, with a per-element axial flag
  // (element_block.cpp:60-87): metric, coefficients, wr, hoop and r per point.
  SetupImage::Block axisym_block(const SetupImage::Domain& d, const std::vector<uint8_t>& axial,
                                 const std::vector<double>& c0, const std::vector<double>& c1,
                                 const std::vector<double>& wr, const std::vector<double>& hoop) const {
    SetupImage::Block b;
    b.kind = INSEM_KIND_AXISYM;
    b.azimuthal_order = am_;
    b.axial = axial;
    const std::size_t ne = d.leaves.size();
    b.elements.resize(ne);
    b.metric.reserve(5 * ne * npe_);
    for (std::size_t e = 0; e < ne; ++e) {
      b.elements[e] = static_cast<int32_t>(e);
      const std::size_t p0 = static_cast<std::size_t>(d.leaves[e]) * npe_;
      for (std::size_t k = 0; k < npe_; ++k) {
        const std::size_t p = p0 + k, q = e * npe_ + k;
        const double m5[5] = {G_.xi_x[p], G_.xi_z[p], G_.eta_x[p], G_.eta_z[p], G_.jac[p]};
        b.metric.insert(b.metric.end(), m5, m5 + 5);
        b.coef0.push_back(static_cast<float>(c0[q]));
        b.coef1.push_back(static_cast<float>(c1[q]));
        b.wr.push_back(static_cast<float>(wr[q]));
        b.hoop.push_back(static_cast<float>(hoop[q]));
        b.r.push_back(static_cast<float>(G_.x[p] > 0.0 ? G_.x[p] : 0.0));
      }
    }
    return b;
  }

  // The elements of one geometry kind, in element order (element_block.cpp:
  // 7-58): Rect keeps (xi_x, eta_z, J) per element, Affine the five metric
  // terms per element, Curved the five per point.
  SetupImage::Block cartesian_block(const SetupImage::Domain& d, Kind kind, const std::vector<double>& c0,
                                    const std::vector<double>& c1) const {
    SetupImage::Block b;
    b.kind = static_cast<int>(kind);
    for (std::size_t e = 0; e < d.leaves.size(); ++e) {
      const int32_t leaf = d.leaves[e];
      if (G_.kind[leaf] != kind) continue;
      b.elements.push_back(static_cast<int32_t>(e));
      const std::size_t p0 = static_cast<std::size_t>(leaf) * npe_;
      const std::size_t points = kind == Kind::Curved ? npe_ : 1;
      for (std::size_t k = 0; k < points; ++k) {
        const std::size_t p = p0 + k;
        if (kind == Kind::Rect) {
          const double m3[3] = {G_.xi_x[p], G_.eta_z[p], G_.jac[p]};
          b.metric.insert(b.metric.end(), m3, m3 + 3);
        } else {
          const double m5[5] = {G_.xi_x[p], G_.xi_z[p], G_.eta_x[p], G_.eta_z[p], G_.jac[p]};
          b.metric.insert(b.metric.end(), m5, m5 + 5);
        }
      }
      for (std::size_t k = 0; k < npe_; ++k) {
        b.coef0.push_back(static_cast<float>(c0[e * npe_ + k]));
        b.coef1.push_back(static_cast<float>(c1[e * npe_ + k]));
      }
    }
    return b;
  }

  // mass = C^T M_e 1: each point's contributions summed in double in
  // ascending slot order, then rounded to float (solver.cpp:230-237).
  void lumped_mass(SetupImage::Domain& d, const std::vector<double>& slot_mass) const {
    const DofMap& m = d.dofs;
    d.mass.resize(static_cast<std::size_t>(m.n_global));
    d.inv_mass.resize(static_cast<std::size_t>(m.n_global));
    for (int32_t g = 0; g < m.n_global; ++g) {
      double sum = 0.0;
      for (int32_t t = m.t_offsets[g]; t < m.t_offsets[g + 1]; ++t) {
        const double v = slot_mass[m.t_slots[t]];
        sum += m.t_weights.empty() ? v : m.t_weights[t] * v;
      }
      INSEM_REQUIRE(sum > 0.0, "point " << g << " of the " << medium_name(d.medium)
                                        << " medium has no positive mass (a degenerate element?)");
      d.mass[g] = static_cast<float>(sum);
      d.inv_mass[g] = static_cast<float>(1.0 / sum);
    }
  }

  // Length element and outward unit normal at element point p on face f,
  // from the stored inverse metric: the face tangent is (x_xi, z_xi) =
  // J (eta_z, -eta_x) along xi and (x_eta, z_eta) = J (-xi_z, xi_x) along
  // eta, counter-clockwise.
  void face_frame(std::size_t p, int f, double& ds, double& nx, double& nz) const {
    const double jac = G_.jac[p];
    const bool along_xi = f == 0 || f == 2;
    double tx = along_xi ? G_.eta_z[p] * jac : -G_.xi_z[p] * jac;
    double tz = along_xi ? -G_.eta_x[p] * jac : G_.xi_x[p] * jac;
    if (f >= 2) tx = -tx, tz = -tz;
    ds = std::hypot(tx, tz);
    nx = tz / ds;
    nz = -tx / ds;
  }

  // Face quadrature weight at face point (i, j) of face f (radius included
  // on axisymmetric meshes) and the outward normal.
  double face_weight(int32_t leaf, int f, int i, int j, double& nx, double& nz) const {
    const std::size_t p = point_of(leaf, i, j);
    double ds;
    face_frame(p, f, ds, nx, nz);
    const bool along_xi = f == 0 || f == 2;
    if (!axisym_) return gll_.weights[along_xi ? i : j] * ds;
    if (along_xi && F_.is_axial(leaf)) {
      const double r_over = i == 0 ? G_.eta_z[p] * G_.jac[p] : G_.x[p] / (1.0 + glj_.points[i]);
      return glj_.weights[i] * ds * r_over;
    }
    return gll_.weights[along_xi ? i : j] * ds * std::max(G_.x[p], 0.0);
  }

  // Boundary conditions of one medium on the macro boundary faces its leaves
  // touch (solver.cpp:256-336): Dirichlet points (elastic rigid, acoustic
  // free) and Stacey damping matrices (absorbing), summed per point over
  // its faces; Dirichlet takes precedence at shared points, and rows and
  // columns of axis components are dropped.
  void boundary_terms(SetupImage::Domain& d) const {
    std::map<int64_t, int> tag_of_face;
    for (const BoundaryFace& bf : M_.boundary()) tag_of_face[static_cast<int64_t>(bf.cell) * 4 + bf.face] = bf.tag;
    const int nc = d.ncomp;
    const bool elastic = d.medium == Medium::Elastic;
    PointSums<9> damping(d.dofs.n_global);
    std::vector<char> dirichlet(static_cast<std::size_t>(d.dofs.n_global), 0);
    for (std::size_t e = 0; e < d.leaves.size(); ++e) {
      const int32_t leaf = d.leaves[e];
      const Leaf& lf = F_.leaf(leaf);
      const int32_t last = (1 << lf.level) - 1;
      const bool on_cell_side[4] = {lf.j == 0, lf.i == last, lf.j == last, lf.i == 0};
      const Material& mat = material(leaf);
      for (int f = 0; f < 4; ++f) {
        if (!on_cell_side[f] || !M_.is_boundary_edge(M_.cell_edge(lf.tree, f))) continue;
        Boundary type = elastic ? Boundary::Free : Boundary::Rigid;  // the natural condition
        const auto tag = tag_of_face.find(static_cast<int64_t>(lf.tree) * 4 + f);
        const bool axis = M_.is_axis_face(lf.tree, f);
        if (tag != tag_of_face.end()) {
          const auto set = boundaries_.find(tag->second);
          INSEM_REQUIRE(!(axis && set != boundaries_.end()),
                        "boundary tag " << tag->second << " lies on the symmetry axis, where no condition applies");
          if (set != boundaries_.end()) type = set->second;
        }
        if (axis) continue;
        const bool clamped = elastic ? type == Boundary::Rigid : type == Boundary::Free;
        if (!clamped && type != Boundary::Absorbing) continue;
        for (int q = 0; q <= n_; ++q) {
          const auto [i, j] = face_point(f, q, n_);
          const int32_t g = d.dofs.global_index(static_cast<int32_t>(e), j * ngll_ + i);
          if (clamped) {
            dirichlet[g] = 1;
            continue;
          }
          double nx, nz;
          const double w = face_weight(leaf, f, i, j, nx, nz);
          stacey_add(damping.at(g), nc, w, mat, nx, nz, elastic);
        }
      }
    }
    std::vector<char> on_axis(static_cast<std::size_t>(d.dofs.n_global), 0);
    for (int32_t g : d.axis_points) on_axis[g] = 1;
    for (int32_t g : damping.ascending()) {
      if (dirichlet[g]) continue;
      double c[9];
      std::copy(damping.get(g), damping.get(g) + 9, c);
      if (on_axis[g])
        for (int a : d.axis_components)
          for (int s = 0; s < nc; ++s) c[a * nc + s] = c[s * nc + a] = 0.0;
      d.absorbing_points.push_back(g);
      for (int s = 0; s < nc * nc; ++s) d.absorbing_coef.push_back(static_cast<float>(c[s]));
    }
    for (int32_t g = 0; g < d.dofs.n_global; ++g)
      if (dirichlet[g]) d.dirichlet_points.push_back(g);
  }

  // Stacey damping of one face point with weight w and normal n
  // (pointwise.hpp:43-57): acoustic w / (rho vp); elastic
  // w rho [vp n n^T + vs (I - n n^T)], for m = 1 turned into (U+, U-, U_z)
  // with the theta direction tangential: T C T^T (solver.cpp:313-319).
  void stacey_add(double* acc, int nc, double w, const Material& m, double nx, double nz, bool elastic) const {
    if (!elastic) {
      acc[0] += w / (m.rho * m.vp);
      return;
    }
    const double p = w * m.rho * m.vp, s = w * m.rho * m.vs;
    const double cxx = p * nx * nx + s * (1.0 - nx * nx);
    const double czz = p * nz * nz + s * (1.0 - nz * nz);
    const double cxz = (p - s) * nx * nz;
    if (nc == 2) {
      const double c4[4] = {cxx, cxz, cxz, czz};
      for (int k = 0; k < 4; ++k) acc[k] += c4[k];
      return;
    }
    const double ctt = w * m.rho * m.vs;
    const double same = 0.5 * (cxx + ctt), cross = 0.5 * (cxx - ctt), rz = kSqrtHalf * cxz;
    const double c9[9] = {same, cross, rz, cross, same, rz, rz, rz, czz};
    for (int k = 0; k < 9; ++k) acc[k] += c9[k];
  }

  // Fluid faces that coincide with solid faces: per fluid point the solid
  // point at the same place and cx, cz = sum over its faces of w n (the
  // fluid's outward normal), in ascending fluid point order
  // (solver.cpp:357-418).
  void coupling() {
    const int fi = img_->domain_index(Medium::Acoustic), si = img_->domain_index(Medium::Elastic);
    if (fi < 0 || si < 0) return;
    const SetupImage::Domain& fluid = img_->domains[fi];
    const SetupImage::Domain& solid = img_->domains[si];
    const Corners cv = F_.corners();
    auto corner = [&](int32_t leaf, int k) { return cv.corner[4 * static_cast<std::size_t>(leaf) + k]; };
    // solid faces by their (lower, higher) vertex pair
    struct SolidFace {
      int32_t lo, hi, element;
      int face;
    };
    std::vector<SolidFace> faces;
    for (std::size_t e = 0; e < solid.leaves.size(); ++e)
      for (int f = 0; f < 4; ++f) {
        const int32_t a = corner(solid.leaves[e], kFaceCorner[f][0]), b = corner(solid.leaves[e], kFaceCorner[f][1]);
        faces.push_back({std::min(a, b), std::max(a, b), static_cast<int32_t>(e), f});
      }
    auto by_pair = [](const SolidFace& x, const SolidFace& y) { return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi; };
    std::sort(faces.begin(), faces.end(), by_pair);
    PointSums<2> normal(fluid.dofs.n_global);
    std::vector<int32_t> partner(static_cast<std::size_t>(fluid.dofs.n_global), -1);
    for (std::size_t e = 0; e < fluid.leaves.size(); ++e) {
      const int32_t leaf = fluid.leaves[e];
      for (int f = 0; f < 4; ++f) {
        const int32_t a = corner(leaf, kFaceCorner[f][0]), b = corner(leaf, kFaceCorner[f][1]);
        const SolidFace key{std::min(a, b), std::max(a, b), 0, 0};
        const auto it = std::lower_bound(faces.begin(), faces.end(), key, by_pair);
        if (it == faces.end() || it->lo != key.lo || it->hi != key.hi) continue;
        const int32_t solid_leaf = solid.leaves[it->element];
        const bool solid_up = corner(solid_leaf, kFaceCorner[it->face][0]) < corner(solid_leaf, kFaceCorner[it->face][1]);
        for (int q = 0; q <= n_; ++q) {
          const auto [i, j] = face_point(f, q, n_);
          const int from_low = a < b ? q : n_ - q;  // position counted from the lower vertex
          const auto [is, js] = face_point(it->face, solid_up ? from_low : n_ - from_low, n_);
          const int32_t gf = fluid.dofs.global_index(static_cast<int32_t>(e), j * ngll_ + i);
          const int32_t gs = solid.dofs.global_index(it->element, js * ngll_ + is);
          INSEM_REQUIRE(partner[gf] < 0 || partner[gf] == gs,
                        "fluid point " << gf << " meets two different solid points on the interface");
          partner[gf] = gs;
          double nx, nz;
          const double w = face_weight(leaf, f, i, j, nx, nz);
          double* c = normal.at(gf);
          c[0] += w * nx;
          c[1] += w * nz;
        }
      }
    }
    for (int32_t gf : normal.ascending()) {
      img_->coupling.fluid_points.push_back(gf);
      img_->coupling.solid_points.push_back(partner[gf]);
      img_->coupling.cx.push_back(static_cast<float>(normal.get(gf)[0]));
      img_->coupling.cz.push_back(static_cast<float>(normal.get(gf)[1]));
    }
  }

  // A point source (solver.cpp:427-520): its weights at every slot of the
  // element that holds it, carried to global points through the slots'
  // constraint rows, scaled by 1/(2 pi) (m = 0) or 1/pi (m = 1) on
  // axisymmetric meshes so that on the axis it is a 3D point source.
  void add_source(std::size_t index, const Source& src) {
    const int di = img_->domain_index(src.medium());
    INSEM_REQUIRE(di >= 0, "source " << index << " (" << src.type_name() << ") acts in the "
                                     << medium_name(src.medium()) << " medium, which this model lacks");
    const SetupImage::Domain& d = img_->domains[di];
    const auto loc = locate(F_, G_, src.position, d.leaves);
    INSEM_REQUIRE(loc, "source " << index << " at (" << src.position[0] << ", " << src.position[1] << ") lies outside the "
                                 << medium_name(src.medium()) << " medium");
    const bool axial = axisym_ && F_.is_axial(loc->leaf);
    const PointWeights w = point_weights(F_, axial ? lag_glj_ : lag_, lag_, *loc);
    const auto& c = src.components;
    const double rs = src.position[0];
    const bool directional = src.type == Source::Force || src.type == Source::Dipole;
    if (axisym_) {
      const bool at_axis = std::fabs(rs) <= 1e-12 * (1.0 + std::fabs(src.position[1]));
      if (am_ == 0) {
        INSEM_REQUIRE(!(at_axis && directional && c[0] != 0.0),
                      "source " << index << " points radially on the axis: that is azimuthal order 1, not 0");
        INSEM_REQUIRE(!(at_axis && src.type == Source::MomentTensor && c[2] != 0.0),
                      "source " << index << " has an M_rz moment on the axis, which azimuthal order 0 cannot hold");
      } else {
        INSEM_REQUIRE(directional, "source " << index << " (" << src.type_name()
                                             << ") is not a force or a dipole, the kinds azimuthal order 1 takes");
        INSEM_REQUIRE(!(at_axis && c[1] != 0.0),
                      "source " << index << " points along the axis on the axis: that is azimuthal order 0, not 1");
      }
    }
    const double scale = !axisym_ ? 1.0 : am_ == 1 ? 1.0 / kPi : 1.0 / (2.0 * kPi);
    const double rho = material(loc->leaf).rho;
    const bool split_radial = d.medium == Medium::Elastic && d.azimuthal_order == 1;
    const int32_t e = d.element_of_leaf[loc->leaf];
    FewPointSums<3> coef;
    for (std::size_t k = 0; k < npe_; ++k) {
      double fc[3] = {0.0, 0.0, 0.0};
      switch (src.type) {
        case Source::Force:  // f phi; m = 1: f_x on U+ and U- by 1/sqrt 2 each
          if (split_radial) {
            fc[0] = fc[1] = kSqrtHalf * c[0] * w.value[k];
            fc[2] = c[1] * w.value[k];
          } else {
            fc[0] = c[0] * w.value[k];
            fc[1] = c[1] * w.value[k];
          }
          break;
        case Source::MomentTensor:  // -div(M delta): M . grad phi, plus M_tt phi / r (M_tt = M_rr)
          fc[0] = c[0] * w.d_dx[k] + c[2] * w.d_dz[k];
          fc[1] = c[2] * w.d_dx[k] + c[1] * w.d_dz[k];
          if (axisym_) fc[0] += c[0] * (rs > 0.0 ? w.value[k] / rs : w.d_dx[k]);
          break;
        case Source::Monopole:
          fc[0] = w.value[k];
          break;
        default:  // dipole -div(f delta / rho): f . grad phi / rho
          fc[0] = (c[0] * w.d_dx[k] + c[1] * w.d_dz[k]) / rho;
          break;
      }
      for (const auto& [g, wt] : d.dofs.slot_entries(e, static_cast<int>(k))) {
        double* acc = coef.at(g);
        for (int comp = 0; comp < d.ncomp; ++comp) acc[comp] += wt * fc[comp] * scale;
      }
    }
    SetupImage::SourceData sd;
    sd.domain = di;
    sd.stf = src.stf;
    sd.points = coef.ascending();
    const std::size_t np = sd.points.size();
    sd.coef.resize(static_cast<std::size_t>(d.ncomp) * np);
    for (std::size_t i = 0; i < np; ++i)
      for (int comp = 0; comp < d.ncomp; ++comp) sd.coef[comp * np + i] = static_cast<float>(coef.get(sd.points[i])[comp]);
    img_->sources.push_back(std::move(sd));
  }

  // A receiver (solver.cpp:522-568): located in the first medium (of those
  // it may record in) that holds it; interpolation and gradient weights per
  // global point through the constraint rows.
  void add_receiver(const Receiver& rec) {
    SetupImage::ReceiverData rd;
    rd.name = rec.name;
    rd.position = rec.position;
    std::optional<Location> loc;
    for (std::size_t di = 0; di < img_->domains.size() && !loc; ++di) {
      if (!rec.medium.empty() && medium_name(img_->domains[di].medium) != rec.medium) continue;
      loc = locate(F_, G_, rec.position, img_->domains[di].leaves);
      if (loc) rd.domain = static_cast<int>(di);
    }
    INSEM_REQUIRE(loc, "receiver '" << rec.name << "' at (" << rec.position[0] << ", " << rec.position[1]
                                    << ") lies outside the " << (rec.medium.empty() ? "model" : rec.medium + " medium"));
    INSEM_REQUIRE(!rec.quantities.empty(), "receiver '" << rec.name << "' has nothing to record");
    const SetupImage::Domain& d = img_->domains[rd.domain];
    const bool axial = axisym_ && F_.is_axial(loc->leaf);
    const PointWeights w = point_weights(F_, axial ? lag_glj_ : lag_, lag_, *loc);
    const int32_t e = d.element_of_leaf[loc->leaf];
    FewPointSums<3> sums;
    for (std::size_t k = 0; k < npe_; ++k)
      for (const auto& [g, wt] : d.dofs.slot_entries(e, static_cast<int>(k))) {
        double* s = sums.at(g);
        s[0] += wt * w.value[k];
        s[1] += wt * w.d_dx[k];
        s[2] += wt * w.d_dz[k];
      }
    for (int32_t g : sums.ascending()) {
      rd.points.push_back(g);
      rd.value.push_back(static_cast<float>(sums.get(g)[0]));
      rd.d_dx.push_back(static_cast<float>(sums.get(g)[1]));
      rd.d_dz.push_back(static_cast<float>(sums.get(g)[2]));
    }
    rd.rho = material(loc->leaf).rho;
    for (const std::string& q : rec.quantities) {
      rd.ncomp.push_back(quantity_ncomp(q, d.medium, d.azimuthal_order));
      rd.quantities.push_back(q);
      rd.codes.push_back(quantity_code(q));
    }
    img_->receivers.push_back(std::move(rd));
  }

  const Forest& F_;
  const Geometry& G_;
  const MacroMesh& M_;
  const Rule& gll_;
  const Rule glj_;
  const Lagrange1D lag_, lag_glj_;
  const std::map<int, Material>& materials_;
  const std::map<int, Boundary>& boundaries_;
  const int ngll_, n_;
  const std::size_t npe_;
  const bool axisym_;
  const int am_;
  std::unique_ptr<SetupImage> img_;
};

}  // namespace

std::unique_ptr<SetupImage> build_setup_image(const Forest& F, const Geometry& G, const Rule& gll, int azimuthal_order,
                                              const std::map<int, Material>& materials,
                                              const std::map<int, Boundary>& boundaries, const TimeConfig& time,
                                              const std::vector<Source>& sources,
                                              const std::vector<Receiver>& receivers) {
  return ImageBuilder(F, G, gll, azimuthal_order, materials, boundaries).build(time, sources, receivers);
}

// ------------------------------------------------------------- setup image

int SetupImage::domain_index(Medium m) const {
  for (std::size_t i = 0; i < domains.size(); ++i)
    if (domains[i].medium == m) return static_cast<int>(i);
  return -1;
}

// Points the C-ABI structs at this image's arrays.
void SetupImage::refresh_views() {
  block_views.assign(domains.size(), {});
  domain_views.assign(domains.size(), insem_hip_domain{});
  for (std::size_t di = 0; di < domains.size(); ++di) {
    const Domain& d = domains[di];
    for (const Block& b : d.blocks) {
      insem_hip_block v{};
      v.kind = b.kind;
      v.azimuthal_order = b.azimuthal_order;
      v.n_elements = static_cast<int32_t>(b.elements.size());
      v.elements = b.elements.data();
      v.metric = b.metric.data();
      v.coef0 = b.coef0.data();
      v.coef1 = b.coef1.data();
      v.wr = b.wr.empty() ? nullptr : b.wr.data();
      v.hoop = b.hoop.empty() ? nullptr : b.hoop.data();
      v.r = b.r.empty() ? nullptr : b.r.data();
      v.axial = b.axial.empty() ? nullptr : b.axial.data();
      block_views[di].push_back(v);
    }
    insem_hip_domain& v = domain_views[di];
    const DofMap& m = d.dofs;
    v.medium = static_cast<int32_t>(d.medium);
    v.ncomp = d.ncomp;
    v.azimuthal_order = d.azimuthal_order;
    v.n_elements = m.n_elements;
    v.n_global = m.n_global;
    v.l2g = m.l2g.data();
    v.n_rows = m.n_rows();
    v.row_offsets = m.row_offsets.data();
    v.row_dofs = m.row_dofs.data();
    v.row_weights = m.row_weights.data();
    v.t_offsets = m.t_offsets.data();
    v.t_slots = m.t_slots.data();
    v.t_weights = m.t_weights.empty() ? nullptr : m.t_weights.data();
    v.mass = d.mass.data();
    v.inv_mass = d.inv_mass.data();
    v.n_absorbing = static_cast<int32_t>(d.absorbing_points.size());
    v.absorbing_points = d.absorbing_points.data();
    v.absorbing_coef = d.absorbing_coef.data();
    v.n_dirichlet = static_cast<int32_t>(d.dirichlet_points.size());
    v.dirichlet_points = d.dirichlet_points.data();
    v.n_axis = static_cast<int32_t>(d.axis_points.size());
    v.axis_points = d.axis_points.data();
    v.n_axis_components = static_cast<int32_t>(d.axis_components.size());
    v.axis_components = d.axis_components.data();
    v.n_blocks = static_cast<int32_t>(block_views[di].size());
    v.blocks = block_views[di].data();
  }
  source_views.clear();
  for (const SourceData& s : sources) source_views.push_back({s.domain, static_cast<int32_t>(s.points.size()),
                                                             s.points.data(), s.coef.data()});
  receiver_views.clear();
  for (const ReceiverData& r : receivers) {
    insem_hip_receiver v{};
    v.domain = r.domain;
    v.n_points = static_cast<int32_t>(r.points.size());
    v.points = r.points.data();
    v.value = r.value.data();
    v.d_dx = r.d_dx.data();
    v.d_dz = r.d_dz.data();
    v.rho = r.rho;
    v.n_quantities = static_cast<int32_t>(r.codes.size());
    v.quantities = r.codes.data();
    receiver_views.push_back(v);
  }
  view = insem_hip_setup{};
  view.ngll = ngll;
  view.D = D.data();
  view.Dj = Dj.data();
  view.w2 = w2.data();
  view.n_domains = static_cast<int32_t>(domain_views.size());
  view.domains = domain_views.data();
  view.coupling = {static_cast<int32_t>(coupling.fluid_points.size()), coupling.fluid_points.data(),
                   coupling.solid_points.data(), coupling.cx.data(), coupling.cz.data()};
  view.n_sources = static_cast<int32_t>(source_views.size());
  view.sources = source_views.data();
  view.n_receivers = static_cast<int32_t>(receiver_views.size());
  view.receivers = receiver_views.data();
  view.record_every = time.record_every;
  view.energy_every = time.energy_every;
}

// Start vectors drawn as the reference draws them (solver.cpp:587-592,
// 629-630): one std::mt19937 seeded 12345 and U(-1, 1) doubles, rounded to
// float; each domain in domain order, then for coupled models the solid's
// and the fluid's again.
namespace {

// The start vectors of every model are a prefix of one stream (one
// generator seeded 12345, drawn in order), kept once per process and
// extended as longer models need it.
class StartStream {
 public:
  std::vector<float> prefix(std::size_t n) {
    std::lock_guard<std::mutex> lock(mu_);
    while (values_.size() < n) values_.push_back(static_cast<float>(u_(gen_)));
    return std::vector<float>(values_.begin(), values_.begin() + static_cast<std::ptrdiff_t>(n));
  }

 private:
  std::mutex mu_;
  std::mt19937 gen_{12345};
  std::uniform_real_distribution<double> u_{-1.0, 1.0};
  std::vector<float> values_;
};

StartStream& start_stream() {
  static StartStream s;
  return s;
}

}  // namespace

std::vector<float> SetupImage::power_start_vectors() const {
  std::size_t n = 0;
  for (const Domain& d : domains) n += static_cast<std::size_t>(d.ncomp) * d.dofs.n_global;
  if (!coupling.fluid_points.empty()) {
    const Domain& s = domains[domain_index(Medium::Elastic)];
    const Domain& f = domains[domain_index(Medium::Acoustic)];
    n += static_cast<std::size_t>(s.ncomp) * s.dofs.n_global + static_cast<std::size_t>(f.ncomp) * f.dofs.n_global;
  }
  return start_stream().prefix(n);
}

std::vector<float> SetupImage::source_amplitudes(double dt, int n_steps) const {
  const std::size_t len = static_cast<std::size_t>(n_steps) + 1;
  std::vector<float> a(sources.size() * len);
  for (std::size_t s = 0; s < sources.size(); ++s)
    for (std::size_t i = 0; i < len; ++i) a[s * len + i] = static_cast<float>(sources[s].stf(static_cast<double>(i) * dt));
  return a;
}

int64_t SetupImage::total_dofs() const {
  int64_t n = 0;
  for (const Domain& d : domains) n += static_cast<int64_t>(d.ncomp) * d.dofs.n_global;
  return n;
}

int64_t SetupImage::total_points() const {
  int64_t n = 0;
  for (const Domain& d : domains) n += d.dofs.n_global;
  return n;
}

}  // namespace insem_amd
