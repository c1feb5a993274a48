// Problem description and the host-side builder of the time-loop state.
//
// The description types carry the reference's user-facing model
// (cpp/include/insem/sim/description.hpp:11-141; Discretization and the
// forest of simulation.hpp:22-48, simulation.cpp:121-185).  SetupImage is
// the state the reference's Solver constructor builds
// (cpp/src/solver/solver.cpp:147-568) -- domains with their point maps,
// element blocks, lumped mass, Stacey/Dirichlet/axis points, fluid-solid
// coupling, sources and receivers -- as owned plain arrays behind the C-ABI
// struct insem_hip_setup (built in setup_image.cpp).
#pragma once

#include <array>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../../../include/insem_hip.h"
#include "basis.hpp"
#include "dofmap.hpp"
#include "geometry.hpp"
#include "mesh.hpp"

namespace insem_amd {

enum class Medium : int { Acoustic = 0, Elastic = 1 };
enum class Boundary : int { Free = 0, Rigid = 1, Absorbing = 2 };

struct Material {
  Medium medium = Medium::Elastic;
  double vp = 0.0, vs = 0.0, rho = 0.0;
  void validate() const;
};

struct Stf {
  enum Type : int { Ricker = 0, Gaussian = 1, GaussianDerivative = 2, Samples = 3 };
  int type = Ricker;
  double f0 = 0.0, t0 = 0.0, amplitude = 1.0;
  std::vector<double> samples;
  double sample_dt = 0.0;
  double operator()(double t) const;
  static Stf make(int type, double f0, double t0, double amplitude);
};

struct Source {
  enum Type : int { Force = 0, MomentTensor = 1, Monopole = 2, Dipole = 3 };
  int type = Force;
  std::array<double, 2> position{0.0, 0.0};
  std::array<double, 3> components{0.0, 0.0, 0.0};
  Stf stf;
  Medium medium() const { return type == Monopole || type == Dipole ? Medium::Acoustic : Medium::Elastic; }
  const char* type_name() const;
};

struct Receiver {
  std::string name;
  std::array<double, 2> position{0.0, 0.0};
  std::vector<std::string> quantities;
  std::string medium;
};

struct TimeConfig {
  double dt = 0.0;
  double safety = 0.8;
  int n_steps = 0;
  double duration = 0.0;
  int record_every = 1;
  int energy_every = 0;
  // Snapshots of whole fields every snapshot_every steps (0: none), of the
  // listed quantities that exist in each medium (displacement, velocity,
  // acceleration in solids; pressure in fluids).
  int snapshot_every = 0;
  std::vector<std::string> snapshot_quantities{"velocity", "pressure"};
};

struct Region {
  double xmin, xmax, zmin, zmax;
  int level;
};

struct Discretization {
  int order = 4;
  int uniform_level = 0;
  double f_max = 0.0;
  double points_per_wavelength = 5.0;
  int max_level = 12;
  std::vector<Region> regions;
  int azimuthal_order = 0;
  bool adaptive() const { return f_max > 0.0 || !regions.empty(); }
};

const char* medium_name(Medium m);
int quantity_ncomp(const std::string& q, Medium m, int azimuthal_order);
int quantity_code(const std::string& q);

// Owning storage behind one insem_hip_setup.
struct SetupImage {
  struct Block {
    int kind = 0;
    int azimuthal_order = 0;
    std::vector<int32_t> elements;
    std::vector<float> metric, coef0, coef1, wr, hoop, r;
    std::vector<uint8_t> axial;
  };
  struct Domain {
    Medium medium = Medium::Elastic;
    int ncomp = 2;
    int azimuthal_order = 0;
    std::vector<int32_t> leaves, element_of_leaf;
    DofMap dofs;
    std::vector<Block> blocks;
    std::vector<float> mass, inv_mass;
    std::vector<int32_t> absorbing_points;
    std::vector<float> absorbing_coef;
    std::vector<int32_t> dirichlet_points, axis_points, axis_components;
    std::vector<double> coords;  // [2][n_global]
    // coordinates of every element point [e * npe + k] (models that take
    // snapshots only)
    std::vector<double> point_x, point_z;
  };
  struct SourceData {
    int domain = 0;
    std::vector<int32_t> points;
    std::vector<float> coef;
    Stf stf;
  };
  struct ReceiverData {
    std::string name;
    std::array<double, 2> position{0.0, 0.0};
    int domain = 0;
    std::vector<int32_t> points;
    std::vector<float> value, d_dx, d_dz;
    double rho = 0.0;
    std::vector<std::string> quantities;
    std::vector<int32_t> codes;
    std::vector<int> ncomp;
  };
  struct CouplingData {
    std::vector<int32_t> fluid_points, solid_points;
    std::vector<float> cx, cz;
  };

  int ngll = 0;
  bool axisym = false;
  int azimuthal_order = 0;
  std::vector<float> D, Dj, w2;
  std::vector<Domain> domains;
  CouplingData coupling;
  std::vector<SourceData> sources;
  std::vector<ReceiverData> receivers;
  TimeConfig time;

  // C-ABI views (valid while this object lives and is not modified).
  insem_hip_setup view{};
  std::vector<insem_hip_domain> domain_views;
  std::vector<std::vector<insem_hip_block>> block_views;
  std::vector<insem_hip_source> source_views;
  std::vector<insem_hip_receiver> receiver_views;
  void refresh_views();

  int domain_index(Medium m) const;
  // Start vectors of the power iteration: mt19937(12345), U(-1, 1), in the
  // reference's draw order (solver.cpp:587-592, 629-630).
  std::vector<float> power_start_vectors() const;
  // amplitudes[s][n] = float(stf_s(n * dt)), n = 0..n_steps.
  std::vector<float> source_amplitudes(double dt, int n_steps) const;
  int64_t total_dofs() const;
  int64_t total_points() const;
};

// The setup image of a set-up model (setup_image.cpp).
std::unique_ptr<SetupImage> build_setup_image(const Forest& F, const Geometry& G, const Rule& gll, int azimuthal_order,
                                              const std::map<int, Material>& materials,
                                              const std::map<int, Boundary>& boundaries, const TimeConfig& time,
                                              const std::vector<Source>& sources,
                                              const std::vector<Receiver>& receivers);

class Model {
 public:
  Model(std::shared_ptr<const MacroMesh> mesh, Discretization disc);

  const MacroMesh& mesh() const { return *mesh_; }
  std::shared_ptr<const MacroMesh> mesh_ptr() const { return mesh_; }
  const Discretization& discretization() const { return disc_; }
  const std::map<int, Material>& materials() const { return materials_; }
  const std::map<int, Boundary>& boundaries() const { return boundaries_; }
  const std::vector<Source>& sources() const { return sources_; }
  const std::vector<Receiver>& receivers() const { return receivers_; }
  void set_material(int region, const Material& m);
  void set_boundary(int tag, Boundary b);
  void add_source(const Source& s);
  void clear_sources();
  void add_receiver(const Receiver& r);
  void clear_receivers();
  void set_time(const TimeConfig& t);
  const TimeConfig& time() const { return time_; }

  void setup();
  bool is_setup() const { return forest_.has_value(); }
  const Forest& forest() const;
  const Geometry& geometry() const;
  const Rule& gll() const;
  // The point map of all leaves (both media), built on first use.
  const DofMap& dofs() const;
  std::map<std::string, int32_t> kind_counts() const;

  // The time-loop state (reference Solver constructor, minus the time step).
  std::unique_ptr<SetupImage> build() const;

 private:
  Forest build_forest() const;
  std::shared_ptr<const MacroMesh> mesh_;
  Discretization disc_;
  std::map<int, Material> materials_;
  std::map<int, Boundary> boundaries_;
  std::vector<Source> sources_;
  std::vector<Receiver> receivers_;
  TimeConfig time_;
  std::optional<Rule> gll_;
  std::optional<Forest> forest_;
  std::optional<Geometry> geom_;
  mutable std::optional<DofMap> dofs_;
};

// Names of the description's enumerations as the reference's files and
// Python interface spell them (cpp/src/sim/description.cpp:17-59,
// cpp/src/mesh/macro_mesh.cpp:54-70); the parsers throw on other names.
const char* boundary_name(Boundary b);
Boundary boundary_from_name(const std::string& s);
Medium medium_from_name(const std::string& s);
const char* stf_name(int type);
const char* insem_version();

}  // namespace insem_amd
