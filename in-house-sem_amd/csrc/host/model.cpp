#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <set>

#include "error.hpp"

namespace insem_amd {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtHalf = 0.70710678118654752440;

// Local (i, j) of point q along face f, counter-clockwise (solver.cpp:28-35).
void face_ij(int f, int q, int n, int& i, int& j) {
  switch (f) {
    case 0: i = q; j = 0; break;
    case 1: i = n; j = q; break;
    case 2: i = n - q; j = n; break;
    default: i = 0; j = n - q; break;
  }
}

// Surface measure and outward normal at a face point from the stored inverse
// metric: x_xi = eta_z J, z_xi = -eta_x J, x_eta = -xi_z J, z_eta = xi_x J
// (solver.cpp:37-60).
void face_measure(const Geometry& g, std::size_t pt, int f, double& ds, double& nx, double& nz) {
  const double jac = g.jac[pt];
  double tx, tz;
  if (f == 0 || f == 2) {
    tx = g.eta_z[pt] * jac;
    tz = -g.eta_x[pt] * jac;
  } else {
    tx = -g.xi_z[pt] * jac;
    tz = g.xi_x[pt] * jac;
  }
  if (f >= 2) {
    tx = -tx;
    tz = -tz;
  }
  ds = std::hypot(tx, tz);
  nx = tz / ds;
  nz = -tx / ds;
}

std::vector<float> to_float(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

}  // namespace

// ------------------------------------------------------------- description

void Material::validate() const {
  INSEM_REQUIRE(rho > 0.0 && std::isfinite(rho), "density must be positive, got " << rho);
  INSEM_REQUIRE(vp > 0.0 && std::isfinite(vp), "P velocity must be positive, got " << vp);
  if (medium == Medium::Elastic) {
    INSEM_REQUIRE(vs > 0.0, "elastic S velocity must be positive, got " << vs << " (use an acoustic material for fluids)");
    INSEM_REQUIRE(vp > vs, "elastic P velocity (" << vp << ") must exceed S velocity (" << vs << ")");
  }
}

Stf Stf::make(int type, double f0, double t0, double amplitude) {
  INSEM_REQUIRE(f0 > 0.0, "source time function frequency must be positive");
  Stf s;
  s.type = type;
  s.f0 = f0;
  s.t0 = t0 < 0.0 ? 1.2 / f0 : t0;
  s.amplitude = amplitude;
  return s;
}

// Source time functions (cpp/src/sim/description.cpp:88-113), in double on the host.
double Stf::operator()(double t) const {
  switch (type) {
    case Ricker: {
      const double a = std::pow(kPi * f0 * (t - t0), 2);
      return amplitude * (1.0 - 2.0 * a) * std::exp(-a);
    }
    case Gaussian: {
      const double a = std::pow(kPi * f0 * (t - t0), 2);
      return amplitude * std::exp(-a);
    }
    case GaussianDerivative: {
      const double tau = t - t0;
      const double c = std::pow(kPi * f0, 2);
      return amplitude * (-2.0 * c * tau) * std::exp(-c * tau * tau);
    }
    case Samples: {
      const double x = t / sample_dt;
      if (x < 0.0 || x > static_cast<double>(samples.size() - 1)) return 0.0;
      const auto i = static_cast<std::size_t>(x);
      if (i + 1 >= samples.size()) return amplitude * samples.back();
      const double w = x - static_cast<double>(i);
      return amplitude * ((1.0 - w) * samples[i] + w * samples[i + 1]);
    }
  }
  return 0.0;
}

const char* Source::type_name() const {
  switch (type) {
    case Force: return "force";
    case MomentTensor: return "moment_tensor";
    case Monopole: return "monopole";
    default: return "dipole";
  }
}

const char* medium_name(Medium m) { return m == Medium::Acoustic ? "acoustic" : "elastic"; }

int quantity_ncomp(const std::string& q, Medium m, int azimuthal_order) {
  if (m == Medium::Elastic) {
    if (q == "displacement" || q == "velocity" || q == "acceleration") return azimuthal_order == 0 ? 2 : 3;
  } else {
    if (q == "pressure") return 1;
    if (q == "particle_velocity") return 2;
  }
  throw Error("quantity '" + q + "' cannot be recorded in " + medium_name(m) +
              " media (elastic: displacement, velocity, acceleration; acoustic: pressure, particle_velocity)");
}

int quantity_code(const std::string& q) {
  if (q == "velocity") return INSEM_Q_VELOCITY;
  if (q == "acceleration") return INSEM_Q_ACCELERATION;
  if (q == "particle_velocity") return INSEM_Q_PARTICLE_VELOCITY;
  return INSEM_Q_FIELD;
}

// ------------------------------------------------------------------- model

Model::Model(std::shared_ptr<const MacroMesh> mesh, Discretization disc) : mesh_(std::move(mesh)), disc_(disc) {
  INSEM_REQUIRE(mesh_, "simulation needs a mesh");
  INSEM_REQUIRE(disc_.order >= 1 && disc_.order <= 16, "polynomial order " << disc_.order << " not in [1, 16]");
  INSEM_REQUIRE(disc_.uniform_level >= 0 && disc_.uniform_level <= Forest::kMaxLevel,
                "uniform refinement level " << disc_.uniform_level << " out of range");
  INSEM_REQUIRE(disc_.max_level >= disc_.uniform_level && disc_.max_level <= Forest::kMaxLevel,
                "max_level " << disc_.max_level << " must be in [uniform_level, " << Forest::kMaxLevel << "]");
  INSEM_REQUIRE(disc_.f_max >= 0.0 && disc_.points_per_wavelength > 0.0, "invalid wavelength refinement parameters");
  INSEM_REQUIRE(disc_.azimuthal_order == 0 || disc_.azimuthal_order == 1,
                "azimuthal order " << disc_.azimuthal_order << " is not supported (0 or 1)");
  INSEM_REQUIRE(disc_.azimuthal_order == 0 || mesh_->axisymmetric(),
                "azimuthal order " << disc_.azimuthal_order << " needs an axisymmetric mesh");
  for (const auto& r : disc_.regions)
    INSEM_REQUIRE(r.xmin <= r.xmax && r.zmin <= r.zmax && r.level >= 0 && r.level <= Forest::kMaxLevel,
                  "invalid refinement region");
}

void Model::set_material(int region, const Material& m) {
  m.validate();
  materials_[region] = m;
  if (disc_.adaptive()) {
    forest_.reset();
    geom_.reset();
  }
}
void Model::set_boundary(int tag, Boundary b) { boundaries_[tag] = b; }
void Model::add_source(const Source& s) { sources_.push_back(s); }
void Model::clear_sources() { sources_.clear(); }
void Model::add_receiver(const Receiver& r) {
  INSEM_REQUIRE(!r.name.empty() && r.name.find('/') == std::string::npos && r.name.find(',') == std::string::npos,
                "receiver name '" << r.name << "' must be non-empty and contain no '/' or ','");
  for (const auto& o : receivers_) INSEM_REQUIRE(o.name != r.name, "duplicate receiver name '" << r.name << "'");
  INSEM_REQUIRE(!r.quantities.empty(), "receiver '" << r.name << "' records no quantities");
  INSEM_REQUIRE(r.medium.empty() || r.medium == "acoustic" || r.medium == "elastic",
                "unknown medium '" << r.medium << "' (expected 'acoustic' or 'elastic')");
  receivers_.push_back(r);
}
void Model::clear_receivers() { receivers_.clear(); }
void Model::set_time(const TimeConfig& t) { time_ = t; }

// Uniform refinement, then wavelength / region refinement, then 2:1 balance
// with equal levels across fluid-solid tree pairs (simulation.cpp:121-175).
Forest Model::build_forest() const {
  Forest forest = Forest::uniform(mesh_, disc_.uniform_level);
  if (disc_.f_max > 0.0)
    for (int region : std::set<int>(mesh_->materials().begin(), mesh_->materials().end()))
      INSEM_REQUIRE(materials_.count(region), "wavelength-based refinement needs a material for region " << region);
  for (;;) {
    std::vector<char> mark(forest.n_leaves(), 0);
    bool any = false;
    for (int32_t l = 0; l < forest.n_leaves(); ++l) {
      const Leaf& lf = forest.leaf(l);
      if (lf.level >= disc_.max_level) continue;
      std::array<std::array<double, 2>, 4> c;
      const double ref[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
      for (int k = 0; k < 4; ++k) c[k] = leaf_map(forest, l, ref[k][0], ref[k][1]);
      bool refine = false;
      if (disc_.f_max > 0.0) {
        double h = 0.0;
        for (int k = 0; k < 4; ++k)
          h = std::max(h, std::hypot(c[(k + 1) % 4][0] - c[k][0], c[(k + 1) % 4][1] - c[k][1]));
        const Material& m = materials_.at(mesh_->material(lf.tree));
        const double vmin = m.medium == Medium::Elastic ? m.vs : m.vp;
        refine = h > disc_.order * (vmin / disc_.f_max) / disc_.points_per_wavelength;
      }
      if (!refine) {
        double x0 = c[0][0], x1 = c[0][0], z0 = c[0][1], z1 = c[0][1];
        for (const auto& p : c) {
          x0 = std::min(x0, p[0]);
          x1 = std::max(x1, p[0]);
          z0 = std::min(z0, p[1]);
          z1 = std::max(z1, p[1]);
        }
        for (const auto& r : disc_.regions)
          if (lf.level < r.level && x0 <= r.xmax && x1 >= r.xmin && z0 <= r.zmax && z1 >= r.zmin) refine = true;
      }
      if (refine) {
        mark[l] = 1;
        any = true;
      }
    }
    if (!any) break;
    forest = forest.refine(mark);
  }
  const auto& mats = materials_;
  const auto mesh = mesh_;
  return forest.balance([&mats, mesh](int32_t a, int32_t b) {
    const auto ma = mats.find(mesh->material(a));
    const auto mb = mats.find(mesh->material(b));
    return ma != mats.end() && mb != mats.end() && ma->second.medium != mb->second.medium;
  });
}

void Model::setup() {
  if (is_setup()) return;
  const int ngll = disc_.order + 1;
  gll_ = gll_rule(ngll);
  forest_ = build_forest();
  geom_ = compute_geometry(*forest_, gll_->points, glj01_rule(ngll).points);
}

const Forest& Model::forest() const {
  INSEM_REQUIRE(is_setup(), "Simulation::setup() has not been called");
  return *forest_;
}
const Geometry& Model::geometry() const {
  INSEM_REQUIRE(is_setup(), "Simulation::setup() has not been called");
  return *geom_;
}
const Rule& Model::gll() const {
  INSEM_REQUIRE(is_setup(), "Simulation::setup() has not been called");
  return *gll_;
}

std::map<std::string, int32_t> Model::kind_counts() const {
  std::map<std::string, int32_t> c{{"rect", 0}, {"affine", 0}, {"curved", 0}};
  for (Kind k : geometry().kind) ++c[k == Kind::Rect ? "rect" : k == Kind::Affine ? "affine" : "curved"];
  return c;
}

std::unique_ptr<SetupImage> Model::build() const {
  INSEM_REQUIRE(is_setup(), "Simulation::setup() has not been called");
  const Forest& F = *forest_;
  const Geometry& G = *geom_;
  const Rule& gll = *gll_;
  const int ngll = gll.size();
  const int n = ngll - 1;
  const std::size_t npe = static_cast<std::size_t>(ngll) * ngll;
  const MacroMesh& M = F.macro();
  const bool axisym = M.axisymmetric();
  const int am = disc_.azimuthal_order;

  INSEM_REQUIRE(am == 0 || (am == 1 && axisym),
                "azimuthal order " << am << " needs an axisymmetric mesh (supported orders: 0, 1)");
  INSEM_REQUIRE(ngll >= 2 && ngll <= 9,
                "polynomial order " << ngll - 1 << " is not supported by the kernels (1..8)");
  INSEM_REQUIRE(time_.record_every >= 1, "record_every must be >= 1");
  INSEM_REQUIRE(time_.energy_every >= 0, "energy_every must be >= 0");

  auto img = std::make_unique<SetupImage>();
  img->ngll = ngll;
  img->axisym = axisym;
  img->azimuthal_order = am;
  img->time = time_;
  const Rule glj = glj01_rule(ngll);
  const Lagrange1D lag(gll.points), lag_glj(glj.points);
  // Kernel tables (kernels.cpp:472-485).
  img->D = to_float(lag.derivative_matrix());
  img->Dj = to_float(lag_glj.derivative_matrix());
  img->w2.resize(npe);
  for (int j = 0; j < ngll; ++j)
    for (int i = 0; i < ngll; ++i) img->w2[j * ngll + i] = static_cast<float>(gll.weights[i] * gll.weights[j]);

  auto material_of_leaf = [&](int32_t leaf) -> const Material& {
    const int region = M.material(F.leaf(leaf).tree);
    return materials_.at(region);
  };
  for (int32_t c = 0; c < M.n_cells(); ++c) {
    const auto it = materials_.find(M.material(c));
    INSEM_REQUIRE(it != materials_.end(), "no material assigned to region " << M.material(c));
    it->second.validate();
  }

  // ---- domains (solver.cpp:147-254)
  for (Medium medium : {Medium::Acoustic, Medium::Elastic}) {
    SetupImage::Domain d;
    d.medium = medium;
    d.azimuthal_order = am;
    d.ncomp = medium == Medium::Elastic ? (am == 1 ? 3 : 2) : 1;
    if (axisym) {
      if (medium == Medium::Elastic)
        d.axis_components = am == 1 ? std::vector<int32_t>{0, 2} : std::vector<int32_t>{0};
      else if (am == 1)
        d.axis_components = {0};
    }
    d.element_of_leaf.assign(F.n_leaves(), -1);
    for (int32_t l = 0; l < F.n_leaves(); ++l)
      if (material_of_leaf(l).medium == medium) {
        d.element_of_leaf[l] = static_cast<int32_t>(d.leaves.size());
        d.leaves.push_back(l);
      }
    if (d.leaves.empty()) continue;
    d.dofs = DofMap::build(F, ngll, d.leaves);

    const std::size_t ne = d.leaves.size();
    const std::size_t nl = ne * npe;
    std::vector<double> wr(nl), hoop(nl, 0.0);
    std::vector<uint8_t> axial(ne, 0);
    for (std::size_t e = 0; e < ne; ++e) {
      const int32_t leaf = d.leaves[e];
      axial[e] = axisym && F.is_axial(leaf) ? 1 : 0;
      for (std::size_t k = 0; k < npe; ++k) {
        const int i = static_cast<int>(k % ngll), j = static_cast<int>(k / ngll);
        const std::size_t pt = static_cast<std::size_t>(leaf) * npe + k;
        const double jac = G.jac[pt];
        const std::size_t q = e * npe + k;
        if (!axisym) {
          wr[q] = gll.weights[i] * gll.weights[j] * jac;
        } else if (!axial[e]) {
          hoop[q] = gll.weights[i] * gll.weights[j] * jac;
          wr[q] = hoop[q] * std::max(G.x[pt], 0.0);
        } else {
          // GLJ along xi: int f r dxi ~ sum_i w^J_i f_i r_i / (1 + xi_i),
          // with dr/dxi as the limit on the axis.
          const double x_xi = G.eta_z[pt] * jac;
          const double rt = i == 0 ? x_xi : G.x[pt] / (1.0 + glj.points[i]);
          wr[q] = glj.weights[i] * gll.weights[j] * jac * rt;
          hoop[q] = i == 0 ? 0.0 : wr[q] / G.x[pt];
        }
      }
    }
    std::vector<double> c0(nl), c1(nl), lmass(nl);
    for (std::size_t e = 0; e < ne; ++e) {
      const Material& m = material_of_leaf(d.leaves[e]);
      for (std::size_t k = 0; k < npe; ++k) {
        const std::size_t q = e * npe + k;
        if (medium == Medium::Elastic) {
          c0[q] = m.rho * (m.vp * m.vp - 2.0 * m.vs * m.vs);
          c1[q] = m.rho * m.vs * m.vs;
          lmass[q] = wr[q] * m.rho;
        } else {
          c0[q] = 1.0 / m.rho;
          c1[q] = 0.0;
          lmass[q] = wr[q] / (m.rho * m.vp * m.vp);
        }
      }
    }
    // Element blocks (element_block.cpp:7-87): Cartesian domains split by
    // kind in the order Rect, Affine, Curved; axisymmetric domains are one
    // Axisym block with a per-element axial flag.
    if (axisym) {
      SetupImage::Block b;
      b.kind = INSEM_KIND_AXISYM;
      b.azimuthal_order = am;
      b.axial = axial;
      for (std::size_t e = 0; e < ne; ++e) {
        b.elements.push_back(static_cast<int32_t>(e));
        const std::size_t p0 = static_cast<std::size_t>(d.leaves[e]) * npe;
        for (std::size_t k = 0; k < npe; ++k) {
          const std::size_t p = p0 + k, q = e * npe + k;
          for (double v : {G.xi_x[p], G.xi_z[p], G.eta_x[p], G.eta_z[p], G.jac[p]}) b.metric.push_back(static_cast<float>(v));
          b.coef0.push_back(static_cast<float>(c0[q]));
          b.coef1.push_back(static_cast<float>(c1[q]));
          b.wr.push_back(static_cast<float>(wr[q]));
          b.hoop.push_back(static_cast<float>(hoop[q]));
          b.r.push_back(static_cast<float>(G.x[p] <= 0.0 ? 0.0 : G.x[p]));
        }
      }
      d.blocks.push_back(std::move(b));
    } else {
      for (Kind kind : {Kind::Rect, Kind::Affine, Kind::Curved}) {
        SetupImage::Block b;
        b.kind = static_cast<int>(kind);
        for (std::size_t e = 0; e < ne; ++e) {
          const int32_t leaf = d.leaves[e];
          if (G.kind[leaf] != kind) continue;
          b.elements.push_back(static_cast<int32_t>(e));
          const std::size_t p0 = static_cast<std::size_t>(leaf) * npe;
          if (kind == Kind::Rect) {
            for (double v : {G.xi_x[p0], G.eta_z[p0], G.jac[p0]}) b.metric.push_back(static_cast<float>(v));
          } else if (kind == Kind::Affine) {
            for (double v : {G.xi_x[p0], G.xi_z[p0], G.eta_x[p0], G.eta_z[p0], G.jac[p0]})
              b.metric.push_back(static_cast<float>(v));
          } else {
            for (std::size_t k = 0; k < npe; ++k) {
              const std::size_t p = p0 + k;
              for (double v : {G.xi_x[p], G.xi_z[p], G.eta_x[p], G.eta_z[p], G.jac[p]})
                b.metric.push_back(static_cast<float>(v));
            }
          }
          for (std::size_t k = 0; k < npe; ++k) {
            b.coef0.push_back(static_cast<float>(c0[e * npe + k]));
            b.coef1.push_back(static_cast<float>(c1[e * npe + k]));
          }
        }
        if (!b.elements.empty()) d.blocks.push_back(std::move(b));
      }
    }
    // Lumped mass C^T M_e 1 in double, then float (solver.cpp:230-237).
    const int32_t ng = d.dofs.n_global;
    std::vector<double> mass(ng, 0.0);
    for (int32_t g = 0; g < ng; ++g) {
      double s = 0.0;
      for (int32_t t = d.dofs.t_offsets[g]; t < d.dofs.t_offsets[g + 1]; ++t) {
        const double v = lmass[d.dofs.t_slots[t]];
        s += d.dofs.t_weights.empty() ? v : d.dofs.t_weights[t] * v;
      }
      mass[g] = s;
    }
    d.mass.assign(mass.begin(), mass.end());
    d.inv_mass.resize(ng);
    for (int32_t g = 0; g < ng; ++g) {
      INSEM_REQUIRE(mass[g] > 0.0, "non-positive lumped mass at a global point (degenerate element?)");
      d.inv_mass[g] = static_cast<float>(1.0 / mass[g]);
    }
    d.coords.resize(2 * static_cast<std::size_t>(ng));
    for (int32_t g = 0; g < ng; ++g) {
      const std::size_t s = static_cast<std::size_t>(d.dofs.primary[g]);
      const std::size_t q = static_cast<std::size_t>(d.leaves[s / npe]) * npe + s % npe;
      d.coords[g] = G.x[q];
      d.coords[ng + g] = G.z[q];
      if (!d.axis_components.empty() && G.x[q] <= 0.0) d.axis_points.push_back(g);
    }
    img->domains.push_back(std::move(d));
  }

  // ---- boundaries (solver.cpp:256-336)
  std::map<int64_t, int> face_tag;
  for (const auto& bf : M.boundary()) face_tag[static_cast<int64_t>(bf.cell) * 4 + bf.face] = bf.tag;
  auto face_weight = [&](int32_t leaf, int f, int i, int j, double& w, double& nx, double& nz) {
    const std::size_t pt = static_cast<std::size_t>(leaf) * npe + static_cast<std::size_t>(j) * ngll + i;
    double ds;
    face_measure(G, pt, f, ds, nx, nz);
    const bool along_xi = f == 0 || f == 2;
    if (!axisym) {
      w = gll.weights[along_xi ? i : j] * ds;
    } else if (along_xi && F.is_axial(leaf)) {
      const double x_xi = G.eta_z[pt] * G.jac[pt];
      const double rt = i == 0 ? x_xi : G.x[pt] / (1.0 + glj.points[i]);
      w = glj.weights[i] * ds * rt;
    } else {
      w = gll.weights[along_xi ? i : j] * ds * std::max(G.x[pt], 0.0);
    }
  };
  for (auto& d : img->domains) {
    const int nc = d.ncomp;
    std::map<int32_t, std::array<double, 9>> absorbing;
    std::set<int32_t> dirichlet;
    for (std::size_t e = 0; e < d.leaves.size(); ++e) {
      const int32_t leaf = d.leaves[e];
      const Leaf& lf = F.leaf(leaf);
      const int last = (1 << lf.level) - 1;
      const bool on_macro[4] = {lf.j == 0, lf.i == last, lf.j == last, lf.i == 0};
      const Material& mat = material_of_leaf(leaf);
      for (int f = 0; f < 4; ++f) {
        if (!on_macro[f] || !M.is_boundary_edge(M.cell_edge(lf.tree, f))) continue;
        Boundary type = d.medium == Medium::Elastic ? Boundary::Free : Boundary::Rigid;
        const auto tag = face_tag.find(static_cast<int64_t>(lf.tree) * 4 + f);
        if (tag != face_tag.end()) {
          const auto bc = boundaries_.find(tag->second);
          if (bc != boundaries_.end()) type = bc->second;
          INSEM_REQUIRE(!(M.is_axis_face(lf.tree, f) && bc != boundaries_.end()),
                        "boundary tag " << tag->second
                                        << " is on the symmetry axis, which takes no boundary condition (remove it "
                                           "from the boundary settings)");
        }
        if (M.is_axis_face(lf.tree, f)) continue;
        const bool is_dir = (d.medium == Medium::Elastic && type == Boundary::Rigid) ||
                            (d.medium == Medium::Acoustic && type == Boundary::Free);
        if (type != Boundary::Absorbing && !is_dir) continue;
        for (int q = 0; q <= n; ++q) {
          int i, j;
          face_ij(f, q, n, i, j);
          const int32_t g = d.dofs.global_index(static_cast<int32_t>(e), j * ngll + i);
          if (is_dir) {
            dirichlet.insert(g);
            continue;
          }
          double w, nx, nz;
          face_weight(leaf, f, i, j, w, nx, nz);
          auto& acc = absorbing[g];  // value-initialised to zero
          if (d.medium == Medium::Acoustic) {
            acc[0] += w / (mat.rho * mat.vp);
            continue;
          }
          // C = w rho [vp n n^T + vs (I - n n^T)] (pointwise.hpp:43-51)
          const double a = w * mat.rho * mat.vp, b = w * mat.rho * mat.vs;
          const double cxx = a * nx * nx + b * (1.0 - nx * nx);
          const double czz = a * nz * nz + b * (1.0 - nz * nz);
          const double cxz = (a - b) * nx * nz;
          if (nc == 2) {
            acc[0] += cxx;
            acc[1] += cxz;
            acc[2] += cxz;
            acc[3] += czz;
          } else {
            // theta tangential; T C T^T into (U+, U-, U_z) (solver.cpp:313-319).
            const double ct = w * mat.rho * mat.vs;
            const double cpp = 0.5 * (cxx + ct), cpm = 0.5 * (cxx - ct), cpz = kSqrtHalf * cxz;
            const double c9[9] = {cpp, cpm, cpz, cpm, cpp, cpz, cpz, cpz, czz};
            for (int s = 0; s < 9; ++s) acc[s] += c9[s];
          }
        }
      }
    }
    const std::set<int32_t> axis(d.axis_points.begin(), d.axis_points.end());
    for (auto [g, c] : absorbing) {
      if (dirichlet.count(g)) continue;
      if (axis.count(g))
        for (int ac : d.axis_components)
          for (int s = 0; s < nc; ++s) c[ac * nc + s] = c[s * nc + ac] = 0.0;
      d.absorbing_points.push_back(g);
      for (int s = 0; s < nc * nc; ++s) d.absorbing_coef.push_back(static_cast<float>(c[s]));
    }
    d.dirichlet_points.assign(dirichlet.begin(), dirichlet.end());
  }

  // ---- fluid-solid coupling (solver.cpp:357-418)
  const int fi = img->domain_index(Medium::Acoustic), si = img->domain_index(Medium::Elastic);
  if (fi >= 0 && si >= 0) {
    const auto& fluid = img->domains[fi];
    const auto& solid = img->domains[si];
    const Corners cv = F.corners();
    auto edge_key = [](int32_t a, int32_t b) {
      return (static_cast<uint64_t>(std::min(a, b)) << 32) | static_cast<uint32_t>(std::max(a, b));
    };
    std::map<uint64_t, std::pair<int32_t, int>> solid_faces;
    for (std::size_t e = 0; e < solid.leaves.size(); ++e) {
      const int32_t* v = &cv.corner[4 * static_cast<std::size_t>(solid.leaves[e])];
      for (int f = 0; f < 4; ++f)
        solid_faces[edge_key(v[kFaceCorner[f][0]], v[kFaceCorner[f][1]])] = {static_cast<int32_t>(e), f};
    }
    struct Entry {
      int32_t solid_point;
      double cx, cz;
    };
    std::map<int32_t, Entry> entries;
    for (std::size_t e = 0; e < fluid.leaves.size(); ++e) {
      const int32_t leaf = fluid.leaves[e];
      const int32_t* v = &cv.corner[4 * static_cast<std::size_t>(leaf)];
      for (int f = 0; f < 4; ++f) {
        const int32_t a = v[kFaceCorner[f][0]], b = v[kFaceCorner[f][1]];
        const auto it = solid_faces.find(edge_key(a, b));
        if (it == solid_faces.end()) continue;
        const auto [es, fs] = it->second;
        const int32_t* vs = &cv.corner[4 * static_cast<std::size_t>(solid.leaves[es])];
        const bool solid_fwd = vs[kFaceCorner[fs][0]] < vs[kFaceCorner[fs][1]];
        for (int q = 0; q <= n; ++q) {
          int i, j;
          face_ij(f, q, n, i, j);
          const int pos = a < b ? q : n - q;
          int is, js;
          face_ij(fs, solid_fwd ? pos : n - pos, n, is, js);
          const int32_t gf = fluid.dofs.global_index(static_cast<int32_t>(e), j * ngll + i);
          const int32_t gs = solid.dofs.global_index(es, js * ngll + is);
          double w, nx, nz;
          face_weight(leaf, f, i, j, w, nx, nz);
          auto ins = entries.emplace(gf, Entry{gs, 0.0, 0.0});
          INSEM_REQUIRE(ins.first->second.solid_point == gs, "inconsistent fluid-solid interface point");
          ins.first->second.cx += w * nx;
          ins.first->second.cz += w * nz;
        }
      }
    }
    for (const auto& [gf, en] : entries) {
      img->coupling.fluid_points.push_back(gf);
      img->coupling.solid_points.push_back(en.solid_point);
      img->coupling.cx.push_back(static_cast<float>(en.cx));
      img->coupling.cz.push_back(static_cast<float>(en.cz));
    }
  }

  // ---- sources (solver.cpp:427-520)
  for (std::size_t sidx = 0; sidx < sources_.size(); ++sidx) {
    const Source& src = sources_[sidx];
    const int di = img->domain_index(src.medium());
    INSEM_REQUIRE(di >= 0, "source " << sidx << " (" << src.type_name() << ") needs an " << medium_name(src.medium())
                                     << " region, but the model has none");
    const auto& d = img->domains[di];
    const auto loc = locate(F, G, src.position, d.leaves);
    INSEM_REQUIRE(loc, "source " << sidx << " at (" << src.position[0] << ", " << src.position[1]
                                 << ") is not inside the " << medium_name(src.medium()) << " part of the model");
    const bool axial = axisym && F.is_axial(loc->leaf);
    const PointWeights w = point_weights(F, axial ? lag_glj : lag, lag, *loc);
    const int32_t e = d.element_of_leaf[loc->leaf];
    const double scale = !axisym ? 1.0 : am == 1 ? 1.0 / kPi : 1.0 / (2.0 * kPi);
    const double rs = src.position[0];
    const auto& c = src.components;
    const bool directional = src.type == Source::Force || src.type == Source::Dipole;
    if (axisym) {
      const bool on_axis = std::fabs(rs) <= 1e-12 * (1.0 + std::fabs(src.position[1]));
      if (am == 0) {
        INSEM_REQUIRE(!(on_axis && directional && c[0] != 0.0),
                      "source " << sidx << ": a radial force or dipole on the axis is not axisymmetric (it is "
                                << "azimuthal order 1)");
        INSEM_REQUIRE(!(on_axis && src.type == Source::MomentTensor && c[2] != 0.0),
                      "source " << sidx << ": an M_rz moment on the axis is not axisymmetric");
      } else {
        INSEM_REQUIRE(directional, "source " << sidx << " (" << src.type_name()
                                             << "): azimuthal order 1 takes force and dipole sources only");
        INSEM_REQUIRE(!(on_axis && c[1] != 0.0), "source " << sidx
                                                           << ": a vertical force or dipole on the axis is "
                                                              "axisymmetric (azimuthal order 0)");
      }
    }
    const double rho = material_of_leaf(loc->leaf).rho;
    const bool pm = d.medium == Medium::Elastic && d.azimuthal_order == 1;
    std::map<int32_t, std::array<double, 3>> acc;
    for (std::size_t k = 0; k < npe; ++k) {
      double fc[3] = {0.0, 0.0, 0.0};
      switch (src.type) {
        case Source::Force:
          if (pm) {
            fc[0] = fc[1] = kSqrtHalf * c[0] * w.value[k];
            fc[2] = c[1] * w.value[k];
          } else {
            fc[0] = c[0] * w.value[k];
            fc[1] = c[1] * w.value[k];
          }
          break;
        case Source::MomentTensor:
          fc[0] = c[0] * w.d_dx[k] + c[2] * w.d_dz[k];
          fc[1] = c[2] * w.d_dx[k] + c[1] * w.d_dz[k];
          if (axisym) fc[0] += c[0] * (rs > 0.0 ? w.value[k] / rs : w.d_dx[k]);
          break;
        case Source::Monopole:
          fc[0] = w.value[k];
          break;
        default:  // Dipole: f . grad(phi)(x_s) / rho
          fc[0] = (c[0] * w.d_dx[k] + c[1] * w.d_dz[k]) / rho;
          break;
      }
      for (const auto& [g, wt] : d.dofs.slot_entries(e, static_cast<int>(k)))
        for (int comp = 0; comp < d.ncomp; ++comp) acc[g][comp] += wt * fc[comp] * scale;
    }
    SetupImage::SourceData sd;
    sd.domain = di;
    sd.stf = src.stf;
    for (const auto& kv : acc) sd.points.push_back(kv.first);
    sd.coef.resize(static_cast<std::size_t>(d.ncomp) * acc.size());
    std::size_t i = 0;
    for (const auto& kv : acc) {
      for (int comp = 0; comp < d.ncomp; ++comp) sd.coef[comp * acc.size() + i] = static_cast<float>(kv.second[comp]);
      ++i;
    }
    img->sources.push_back(std::move(sd));
  }

  // ---- receivers (solver.cpp:522-568)
  for (const Receiver& rec : receivers_) {
    SetupImage::ReceiverData rd;
    rd.name = rec.name;
    rd.position = rec.position;
    std::optional<Location> loc;
    for (std::size_t di = 0; di < img->domains.size() && !loc; ++di) {
      if (!rec.medium.empty() && medium_name(img->domains[di].medium) != rec.medium) continue;
      loc = locate(F, G, rec.position, img->domains[di].leaves);
      if (loc) rd.domain = static_cast<int>(di);
    }
    INSEM_REQUIRE(loc, "receiver '" << rec.name << "' at (" << rec.position[0] << ", " << rec.position[1]
                                    << ") is outside the " << (rec.medium.empty() ? "" : rec.medium + " part of the ")
                                    << "model");
    const auto& d = img->domains[rd.domain];
    const bool axial = axisym && F.is_axial(loc->leaf);
    const PointWeights w = point_weights(F, axial ? lag_glj : lag, lag, *loc);
    const int32_t e = d.element_of_leaf[loc->leaf];
    std::map<int32_t, std::array<double, 3>> acc;
    for (std::size_t k = 0; k < npe; ++k)
      for (const auto& [g, wt] : d.dofs.slot_entries(e, static_cast<int>(k))) {
        acc[g][0] += wt * w.value[k];
        acc[g][1] += wt * w.d_dx[k];
        acc[g][2] += wt * w.d_dz[k];
      }
    for (const auto& [g, c] : acc) {
      rd.points.push_back(g);
      rd.value.push_back(static_cast<float>(c[0]));
      rd.d_dx.push_back(static_cast<float>(c[1]));
      rd.d_dz.push_back(static_cast<float>(c[2]));
    }
    rd.rho = material_of_leaf(loc->leaf).rho;
    INSEM_REQUIRE(!rec.quantities.empty(), "receiver '" << rec.name << "' records no quantities");
    for (const std::string& q : rec.quantities) {
      rd.ncomp.push_back(quantity_ncomp(q, d.medium, d.azimuthal_order));
      rd.quantities.push_back(q);
      rd.codes.push_back(quantity_code(q));
    }
    img->receivers.push_back(std::move(rd));
  }
  img->refresh_views();
  return img;
}

// ------------------------------------------------------------- setup image

int SetupImage::domain_index(Medium m) const {
  for (std::size_t i = 0; i < domains.size(); ++i)
    if (domains[i].medium == m) return static_cast<int>(i);
  return -1;
}

void SetupImage::refresh_views() {
  domain_views.clear();
  block_views.assign(domains.size(), {});
  for (std::size_t di = 0; di < domains.size(); ++di) {
    const Domain& d = domains[di];
    for (const Block& b : d.blocks) {
      insem_hip_block bv{};
      bv.kind = b.kind;
      bv.azimuthal_order = b.azimuthal_order;
      bv.n_elements = static_cast<int32_t>(b.elements.size());
      bv.elements = b.elements.data();
      bv.metric = b.metric.data();
      bv.coef0 = b.coef0.data();
      bv.coef1 = b.coef1.data();
      bv.wr = b.wr.empty() ? nullptr : b.wr.data();
      bv.hoop = b.hoop.empty() ? nullptr : b.hoop.data();
      bv.r = b.r.empty() ? nullptr : b.r.data();
      bv.axial = b.axial.empty() ? nullptr : b.axial.data();
      block_views[di].push_back(bv);
    }
    insem_hip_domain dv{};
    dv.medium = static_cast<int32_t>(d.medium);
    dv.ncomp = d.ncomp;
    dv.azimuthal_order = d.azimuthal_order;
    dv.n_elements = d.dofs.n_elements;
    dv.n_global = d.dofs.n_global;
    dv.l2g = d.dofs.l2g.data();
    dv.n_rows = d.dofs.n_rows();
    dv.row_offsets = d.dofs.row_offsets.data();
    dv.row_dofs = d.dofs.row_dofs.data();
    dv.row_weights = d.dofs.row_weights.data();
    dv.t_offsets = d.dofs.t_offsets.data();
    dv.t_slots = d.dofs.t_slots.data();
    dv.t_weights = d.dofs.t_weights.empty() ? nullptr : d.dofs.t_weights.data();
    dv.mass = d.mass.data();
    dv.inv_mass = d.inv_mass.data();
    dv.n_absorbing = static_cast<int32_t>(d.absorbing_points.size());
    dv.absorbing_points = d.absorbing_points.data();
    dv.absorbing_coef = d.absorbing_coef.data();
    dv.n_dirichlet = static_cast<int32_t>(d.dirichlet_points.size());
    dv.dirichlet_points = d.dirichlet_points.data();
    dv.n_axis = static_cast<int32_t>(d.axis_points.size());
    dv.axis_points = d.axis_points.data();
    dv.n_axis_components = static_cast<int32_t>(d.axis_components.size());
    dv.axis_components = d.axis_components.data();
    dv.n_blocks = static_cast<int32_t>(block_views[di].size());
    dv.blocks = block_views[di].data();
    domain_views.push_back(dv);
  }
  source_views.clear();
  for (const auto& s : sources) {
    insem_hip_source sv{};
    sv.domain = s.domain;
    sv.n_points = static_cast<int32_t>(s.points.size());
    sv.points = s.points.data();
    sv.coef = s.coef.data();
    source_views.push_back(sv);
  }
  receiver_views.clear();
  for (const auto& r : receivers) {
    insem_hip_receiver rv{};
    rv.domain = r.domain;
    rv.n_points = static_cast<int32_t>(r.points.size());
    rv.points = r.points.data();
    rv.value = r.value.data();
    rv.d_dx = r.d_dx.data();
    rv.d_dz = r.d_dz.data();
    rv.rho = r.rho;
    rv.n_quantities = static_cast<int32_t>(r.codes.size());
    rv.quantities = r.codes.data();
    receiver_views.push_back(rv);
  }
  view = insem_hip_setup{};
  view.ngll = ngll;
  view.D = D.data();
  view.Dj = Dj.data();
  view.w2 = w2.data();
  view.n_domains = static_cast<int32_t>(domain_views.size());
  view.domains = domain_views.data();
  view.coupling.n_points = static_cast<int32_t>(coupling.fluid_points.size());
  view.coupling.fluid_points = coupling.fluid_points.data();
  view.coupling.solid_points = coupling.solid_points.data();
  view.coupling.cx = coupling.cx.data();
  view.coupling.cz = coupling.cz.data();
  view.n_sources = static_cast<int32_t>(source_views.size());
  view.sources = source_views.data();
  view.n_receivers = static_cast<int32_t>(receiver_views.size());
  view.receivers = receiver_views.data();
  view.record_every = time.record_every;
  view.energy_every = time.energy_every;
}

std::vector<float> SetupImage::power_start_vectors() const {
  std::mt19937 rng(12345);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  std::vector<float> out;
  for (const auto& d : domains) {
    const std::size_t len = static_cast<std::size_t>(d.ncomp) * d.dofs.n_global;
    for (std::size_t i = 0; i < len; ++i) out.push_back(static_cast<float>(dist(rng)));
  }
  if (!coupling.fluid_points.empty()) {
    for (Medium m : {Medium::Elastic, Medium::Acoustic}) {
      const auto& d = domains[domain_index(m)];
      const std::size_t len = static_cast<std::size_t>(d.ncomp) * d.dofs.n_global;
      for (std::size_t i = 0; i < len; ++i) out.push_back(static_cast<float>(dist(rng)));
    }
  }
  return out;
}

std::vector<float> SetupImage::source_amplitudes(double dt, int n_steps) const {
  std::vector<float> a(sources.size() * (static_cast<std::size_t>(n_steps) + 1));
  for (std::size_t s = 0; s < sources.size(); ++s)
    for (int i = 0; i <= n_steps; ++i)
      a[s * (static_cast<std::size_t>(n_steps) + 1) + i] = static_cast<float>(sources[s].stf(i * dt));
  return a;
}

int64_t SetupImage::total_dofs() const {
  int64_t s = 0;
  for (const auto& d : domains) s += static_cast<int64_t>(d.ncomp) * d.dofs.n_global;
  return s;
}

int64_t SetupImage::total_points() const {
  int64_t s = 0;
  for (const auto& d : domains) s += d.dofs.n_global;
  return s;
}

}  // namespace insem_amd
