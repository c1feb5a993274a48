#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <set>

#include "error.hpp"

namespace insem_amd {
namespace {

constexpr double kPi = 3.14159265358979323846;

}  // namespace

// ------------------------------------------------------------- description

void Material::validate() const {
  INSEM_REQUIRE(std::isfinite(rho) && rho > 0.0, "a material needs a positive density, not " << rho);
  INSEM_REQUIRE(std::isfinite(vp) && vp > 0.0, "a material needs a positive P velocity, not " << vp);
  if (medium == Medium::Elastic) {
    INSEM_REQUIRE(vs > 0.0, "an elastic material needs a positive S velocity, not " << vs
                                                                                  << " (fluids are acoustic materials)");
    INSEM_REQUIRE(vs < vp, "the S velocity " << vs << " must be below the P velocity " << vp);
  }
}

Stf Stf::make(int type, double f0, double t0, double amplitude) {
  INSEM_REQUIRE(f0 > 0.0, "a source time function needs a positive frequency, not " << f0);
  Stf s;
  s.type = type;
  s.f0 = f0;
  s.t0 = t0 < 0.0 ? 1.2 / f0 : t0;  // default delay 1.2 / f0 (description.cpp:13)
  s.amplitude = amplitude;
  return s;
}

// The source time functions of cpp/src/sim/description.cpp:88-113, in
// double: Ricker (1 - 2a) e^-a and Gaussian e^-a with a = (pi f0 (t - t0))^2,
// the Gaussian's derivative, or samples interpolated linearly (0 outside).
double Stf::operator()(double t) const {
  if (type == Samples) {
    const double x = t / sample_dt;
    if (x < 0.0 || x > static_cast<double>(samples.size() - 1)) return 0.0;
    const auto i = static_cast<std::size_t>(x);
    if (i + 1 >= samples.size()) return amplitude * samples.back();
    const double frac = x - static_cast<double>(i);
    return amplitude * ((1.0 - frac) * samples[i] + frac * samples[i + 1]);
  }
  const double tau = t - t0;
  if (type == GaussianDerivative) {
    const double c = std::pow(kPi * f0, 2);
    return amplitude * (-2.0 * c * tau) * std::exp(-c * tau * tau);
  }
  const double a = std::pow(kPi * f0 * tau, 2);
  return type == Ricker ? amplitude * (1.0 - 2.0 * a) * std::exp(-a) : amplitude * std::exp(-a);
}

const char* Source::type_name() const {
  static const char* names[] = {"force", "moment_tensor", "monopole", "dipole"};
  return type >= 0 && type < 4 ? names[type] : "dipole";
}

const char* medium_name(Medium m) { return m == Medium::Acoustic ? "acoustic" : "elastic"; }

const char* boundary_name(Boundary b) {
  return b == Boundary::Free ? "free" : b == Boundary::Rigid ? "rigid" : "absorbing";
}

Boundary boundary_from_name(const std::string& s) {
  for (Boundary b : {Boundary::Free, Boundary::Rigid, Boundary::Absorbing})
    if (s == boundary_name(b)) return b;
  throw Error("boundary type '" + s + "' is none of 'free', 'rigid', 'absorbing'");
}

Medium medium_from_name(const std::string& s) {
  if (s == "acoustic") return Medium::Acoustic;
  if (s == "elastic") return Medium::Elastic;
  throw Error("medium '" + s + "' is neither 'acoustic' nor 'elastic'");
}

const char* stf_name(int type) {
  static const char* names[] = {"ricker", "gaussian", "gaussian_derivative", "samples"};
  return type >= 0 && type < 4 ? names[type] : "samples";
}

// The version this backend reports as insem's (the reference project's
// version, CMakeLists.txt:3), so that files and front ends agree on it.
const char* insem_version() { return "0.1.0"; }

// Components of a recorded quantity (description.hpp:225-244): elastic
// displacement, velocity and acceleration have 2 (m = 0) or 3 (m = 1:
// U_r, U_theta, U_z); acoustic pressure 1 and particle velocity 2.
int quantity_ncomp(const std::string& q, Medium m, int azimuthal_order) {
  const bool elastic_field = q == "displacement" || q == "velocity" || q == "acceleration";
  if (m == Medium::Elastic && elastic_field) return azimuthal_order == 0 ? 2 : 3;
  if (m == Medium::Acoustic && q == "pressure") return 1;
  if (m == Medium::Acoustic && q == "particle_velocity") return 2;
  throw Error("'" + q + "' is not recorded in " + medium_name(m) +
              " media, which record " +
              (m == Medium::Elastic ? "displacement, velocity and acceleration" : "pressure and particle_velocity"));
}

int quantity_code(const std::string& q) {
  if (q == "velocity") return INSEM_Q_VELOCITY;
  if (q == "acceleration") return INSEM_Q_ACCELERATION;
  if (q == "particle_velocity") return INSEM_Q_PARTICLE_VELOCITY;
  return INSEM_Q_FIELD;
}

// ------------------------------------------------------------------- model

Model::Model(std::shared_ptr<const MacroMesh> mesh, Discretization disc) : mesh_(std::move(mesh)), disc_(disc) {
  INSEM_REQUIRE(mesh_, "a model needs a mesh");
  INSEM_REQUIRE(disc_.order >= 1 && disc_.order <= 16, "the polynomial order must be 1..16, not " << disc_.order);
  INSEM_REQUIRE(disc_.uniform_level >= 0 && disc_.uniform_level <= disc_.max_level && disc_.max_level <= Forest::kMaxLevel,
                "refinement levels must satisfy 0 <= uniform_level (" << disc_.uniform_level << ") <= max_level ("
                                                                      << disc_.max_level << ") <= " << Forest::kMaxLevel);
  INSEM_REQUIRE(disc_.f_max >= 0.0 && disc_.points_per_wavelength > 0.0,
                "wavelength refinement needs f_max >= 0 and points_per_wavelength > 0");
  INSEM_REQUIRE(disc_.azimuthal_order == 0 || disc_.azimuthal_order == 1,
                "the azimuthal order is 0 or 1, not " << disc_.azimuthal_order);
  INSEM_REQUIRE(disc_.azimuthal_order == 0 || mesh_->axisymmetric(),
                "azimuthal order " << disc_.azimuthal_order << " needs an axisymmetric mesh");
  for (const Region& r : disc_.regions)
    INSEM_REQUIRE(r.xmin <= r.xmax && r.zmin <= r.zmax && r.level >= 0 && r.level <= Forest::kMaxLevel,
                  "a refinement region needs xmin <= xmax, zmin <= zmax and a level in 0.." << Forest::kMaxLevel);
}

void Model::set_material(int region, const Material& m) {
  m.validate();
  materials_[region] = m;
  if (disc_.adaptive()) {  // wavelength refinement depends on the materials
    forest_.reset();
    geom_.reset();
    dofs_.reset();
  }
}
void Model::set_boundary(int tag, Boundary b) { boundaries_[tag] = b; }
void Model::add_source(const Source& s) { sources_.push_back(s); }
void Model::clear_sources() { sources_.clear(); }
void Model::add_receiver(const Receiver& r) {
  INSEM_REQUIRE(!r.name.empty() && r.name.find_first_of("/,") == std::string::npos,
                "receiver names are non-empty and hold no '/' or ',': '" << r.name << "'");
  for (const Receiver& o : receivers_) INSEM_REQUIRE(o.name != r.name, "a receiver named '" << r.name << "' exists");
  INSEM_REQUIRE(!r.quantities.empty(), "receiver '" << r.name << "' has nothing to record");
  INSEM_REQUIRE(r.medium.empty() || r.medium == "acoustic" || r.medium == "elastic",
                "receiver '" << r.name << "' names medium '" << r.medium << "'; it is 'acoustic' or 'elastic'");
  receivers_.push_back(r);
}
void Model::clear_receivers() { receivers_.clear(); }
void Model::set_time(const TimeConfig& t) { time_ = t; }

namespace {

// Longest side of a leaf's corner quadrilateral, and whether its corner box
// meets a refinement region it is coarser than.
double longest_side(const std::array<std::array<double, 2>, 4>& c) {
  double h = 0.0;
  for (int k = 0; k < 4; ++k) h = std::max(h, std::hypot(c[(k + 1) % 4][0] - c[k][0], c[(k + 1) % 4][1] - c[k][1]));
  return h;
}

bool meets_region(const std::array<std::array<double, 2>, 4>& c, int level, const std::vector<Region>& regions) {
  double x0 = c[0][0], x1 = x0, z0 = c[0][1], z1 = z0;
  for (const auto& p : c) {
    x0 = std::min(x0, p[0]), x1 = std::max(x1, p[0]);
    z0 = std::min(z0, p[1]), z1 = std::max(z1, p[1]);
  }
  for (const Region& r : regions)
    if (level < r.level && x0 <= r.xmax && x1 >= r.xmin && z0 <= r.zmax && z1 >= r.zmin) return true;
  return false;
}

}  // namespace

// The forest of simulation.cpp:121-175: the uniform level, then rounds of
// refinement of every leaf (below max_level) whose longest side exceeds
// order * wavelength / points_per_wavelength at f_max (the S wavelength in
// solids, P in fluids) or that meets a refinement region it is coarser
// than, until none is left; then 2:1 balance, with fluid and solid leaves
// meeting at equal levels.
Forest Model::build_forest() const {
  if (disc_.f_max > 0.0)
    for (int region : std::set<int>(mesh_->materials().begin(), mesh_->materials().end()))
      INSEM_REQUIRE(materials_.count(region), "wavelength refinement needs the material of region " << region);
  Forest forest = Forest::uniform(mesh_, disc_.uniform_level);
  constexpr double kCorner[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  for (;;) {
    std::vector<char> mark(static_cast<std::size_t>(forest.n_leaves()), 0);
    bool marked = false;
    for (int32_t l = 0; l < forest.n_leaves(); ++l) {
      const Leaf& lf = forest.leaf(l);
      if (lf.level >= disc_.max_level) continue;
      std::array<std::array<double, 2>, 4> c;
      for (int k = 0; k < 4; ++k) c[k] = leaf_map(forest, l, kCorner[k][0], kCorner[k][1]);
      bool split = false;
      if (disc_.f_max > 0.0) {
        const Material& m = materials_.at(mesh_->material(lf.tree));
        const double slowest = m.medium == Medium::Elastic ? m.vs : m.vp;
        split = longest_side(c) > disc_.order * (slowest / disc_.f_max) / disc_.points_per_wavelength;
      }
      if (!split) split = meets_region(c, lf.level, disc_.regions);
      if (split) mark[l] = marked = 1;
    }
    if (!marked) break;
    forest = forest.refine(mark);
  }
  const std::map<int, Material>& mats = materials_;
  const std::shared_ptr<const MacroMesh> mesh = mesh_;
  return forest.balance([&mats, mesh](int32_t a, int32_t b) {
    const auto ma = mats.find(mesh->material(a)), mb = mats.find(mesh->material(b));
    return ma != mats.end() && mb != mats.end() && ma->second.medium != mb->second.medium;
  });
}

void Model::setup() {
  if (is_setup()) return;
  const int ngll = disc_.order + 1;
  gll_ = gll_rule(ngll);
  forest_ = build_forest();
  dofs_.reset();
  geom_ = compute_geometry(*forest_, gll_->points, glj01_rule(ngll).points);
}

const Forest& Model::forest() const {
  INSEM_REQUIRE(is_setup(), "the model is not set up yet (call setup first)");
  return *forest_;
}
const Geometry& Model::geometry() const {
  INSEM_REQUIRE(is_setup(), "the model is not set up yet (call setup first)");
  return *geom_;
}
const DofMap& Model::dofs() const {
  if (!dofs_) {
    std::vector<int32_t> all(static_cast<std::size_t>(forest().n_leaves()));
    for (int32_t l = 0; l < forest().n_leaves(); ++l) all[l] = l;
    dofs_ = DofMap::build(forest(), disc_.order + 1, all);
  }
  return *dofs_;
}

const Rule& Model::gll() const {
  INSEM_REQUIRE(is_setup(), "the model is not set up yet (call setup first)");
  return *gll_;
}

std::map<std::string, int32_t> Model::kind_counts() const {
  std::map<std::string, int32_t> c{{"rect", 0}, {"affine", 0}, {"curved", 0}};
  static const char* names[] = {"rect", "affine", "curved", "curved"};
  for (Kind k : geometry().kind) ++c[names[static_cast<int>(k)]];
  return c;
}

std::unique_ptr<SetupImage> Model::build() const {
  INSEM_REQUIRE(is_setup(), "the model is not set up yet (call setup first)");
  return build_setup_image(*forest_, *geom_, *gll_, disc_.azimuthal_order, materials_, boundaries_, time_, sources_,
                           receivers_);
}

}  // namespace insem_amd
