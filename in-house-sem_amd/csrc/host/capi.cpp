// Flat C API of the host library for the Python package (the role the
// reference's pybind11 module plays, python/src/bindings.cpp:110-494).
// Handles are opaque pointers; every call returns 0 or -1 and leaves the
// message of a failure in insem_host_last_error().
#include <cstdint>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/insem_hip.h"
#include "basis.hpp"
#include "error.hpp"
#include "model.hpp"
#include "batch.hpp"
#include "solver.hpp"

using namespace insem_amd;

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

template <class F>
void* guard_ptr(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

struct MeshHandle {
  std::shared_ptr<const MacroMesh> mesh;
};

// The object behind a handle; a null handle (a closed object, or an image a
// batch or solver took over) is an error, not a crash.
template <class T>
T* ref(void* h) {
  INSEM_REQUIRE(h, "null handle: the object was closed or taken over");
  return static_cast<T*>(h);
}

std::vector<std::string> split_csv(const char* s) {
  std::vector<std::string> out;
  std::stringstream ss(s ? s : "");
  std::string item;
  while (std::getline(ss, item, ','))
    if (!item.empty()) out.push_back(item);
  return out;
}

}  // namespace

extern "C" {

const char* insem_host_last_error(void) { return g_err.c_str(); }

// ----------------------------------------------------------------- basis

int insem_basis_rule(int32_t n, int32_t alpha, int32_t beta, double* points, double* weights) {
  return guard([&] {
    const Rule r = lobatto_rule(n, alpha, beta);
    std::memcpy(points, r.points.data(), n * sizeof(double));
    std::memcpy(weights, r.weights.data(), n * sizeof(double));
  });
}

int insem_basis_derivative_matrix(int32_t n, const double* nodes, double* out) {
  return guard([&] {
    const auto d = Lagrange1D(std::vector<double>(nodes, nodes + n)).derivative_matrix();
    std::memcpy(out, d.data(), d.size() * sizeof(double));
  });
}

// ------------------------------------------------------------------ mesh

void* insem_mesh_new(int32_t npc, int32_t axisym, const double* nodes, int32_t n_nodes, const int32_t* cells,
                     int32_t n_cells, const int32_t* material, const int32_t* boundary, int32_t n_boundary) {
  return guard_ptr([&]() -> void* {
    std::vector<BoundaryFace> bf;
    for (int i = 0; i < n_boundary; ++i) bf.push_back({boundary[3 * i], boundary[3 * i + 1], boundary[3 * i + 2]});
    auto m = std::make_shared<const MacroMesh>(
        npc, axisym != 0, std::vector<double>(nodes, nodes + 2 * static_cast<size_t>(n_nodes)),
        std::vector<int32_t>(cells, cells + static_cast<size_t>(n_cells) * npc),
        std::vector<int>(material, material + n_cells), std::move(bf));
    return new MeshHandle{m};
  });
}

void insem_mesh_free(void* h) { delete static_cast<MeshHandle*>(h); }

int insem_mesh_info(void* h, int32_t* out) {
  return guard([&] {
    const MacroMesh& m = *ref<MeshHandle>(h)->mesh;
    out[0] = m.n_nodes();
    out[1] = m.n_cells();
    out[2] = m.n_edges();
    out[3] = m.nodes_per_cell();
    out[4] = m.axisymmetric() ? 1 : 0;
    out[5] = static_cast<int32_t>(m.boundary().size());
  });
}

int insem_mesh_arrays(void* h, int32_t* cells, int32_t* boundary) {
  return guard([&] {
    const MacroMesh& m = *ref<MeshHandle>(h)->mesh;
    if (cells) std::memcpy(cells, m.cells().data(), m.cells().size() * sizeof(int32_t));
    if (boundary)
      for (size_t i = 0; i < m.boundary().size(); ++i) {
        boundary[3 * i] = m.boundary()[i].cell;
        boundary[3 * i + 1] = m.boundary()[i].face;
        boundary[3 * i + 2] = m.boundary()[i].tag;
      }
  });
}

int insem_mesh_map(void* h, int32_t cell, double xi, double eta, double* out) {
  return guard([&] {
    const MacroMesh& m = *ref<MeshHandle>(h)->mesh;
    INSEM_REQUIRE(cell >= 0 && cell < m.n_cells(), "cell " << cell << " out of range");
    const auto x = m.map(cell, xi, eta);
    out[0] = x[0];
    out[1] = x[1];
  });
}

int insem_mesh_is_axis_face(void* h, int32_t cell, int32_t face) {
  int out = -1;
  (void)guard([&] { out = ref<MeshHandle>(h)->mesh->is_axis_face(cell, face) ? 1 : 0; });
  return out;
}

// ----------------------------------------------------------------- model

void* insem_model_new(void* mesh, int32_t order, int32_t uniform_level, double f_max, double ppw, int32_t max_level,
                      const double* regions, int32_t n_regions, int32_t azimuthal_order) {
  return guard_ptr([&]() -> void* {
    Discretization d;
    d.order = order;
    d.uniform_level = uniform_level;
    d.f_max = f_max;
    d.points_per_wavelength = ppw;
    d.max_level = max_level;
    d.azimuthal_order = azimuthal_order;
    for (int i = 0; i < n_regions; ++i)
      d.regions.push_back({regions[5 * i], regions[5 * i + 1], regions[5 * i + 2], regions[5 * i + 3],
                           static_cast<int>(regions[5 * i + 4])});
    return new Model(static_cast<MeshHandle*>(mesh)->mesh, d);
  });
}

void insem_model_free(void* h) { delete static_cast<Model*>(h); }

int insem_model_set_material(void* h, int32_t region, int32_t medium, double vp, double vs, double rho) {
  return guard([&] {
    Material m;
    m.medium = medium == 0 ? Medium::Acoustic : Medium::Elastic;
    m.vp = vp;
    m.vs = medium == 0 ? 0.0 : vs;
    m.rho = rho;
    ref<Model>(h)->set_material(region, m);
  });
}

int insem_model_set_boundary(void* h, int32_t tag, int32_t type) {
  return guard([&] {
    INSEM_REQUIRE(type >= 0 && type <= 2, "unknown boundary type " << type);
    ref<Model>(h)->set_boundary(tag, static_cast<Boundary>(type));
  });
}

int insem_model_add_source(void* h, int32_t type, double x, double z, double c0, double c1, double c2, int32_t stf_type,
                           double f0, double t0, double amp, const double* samples, int32_t n_samples,
                           double sample_dt) {
  return guard([&] {
    INSEM_REQUIRE(type >= 0 && type <= 3, "unknown source type " << type);
    Source s;
    s.type = type;
    s.position = {x, z};
    s.components = {c0, c1, c2};
    if (stf_type == Stf::Samples) {
      INSEM_REQUIRE(sample_dt > 0.0, "sample interval must be positive");
      INSEM_REQUIRE(n_samples > 0, "sampled source time function needs samples");
      s.stf.type = Stf::Samples;
      s.stf.samples.assign(samples, samples + n_samples);
      s.stf.sample_dt = sample_dt;
      s.stf.amplitude = amp;
    } else {
      s.stf = Stf::make(stf_type, f0, t0, amp);
    }
    ref<Model>(h)->add_source(s);
  });
}

int insem_model_clear_sources(void* h) {
  return guard([&] { ref<Model>(h)->clear_sources(); });
}

int insem_model_add_receiver(void* h, const char* name, double x, double z, const char* quantities, const char* medium) {
  return guard([&] {
    Receiver r;
    r.name = name ? name : "";
    r.position = {x, z};
    r.quantities = split_csv(quantities);
    r.medium = medium ? medium : "";
    ref<Model>(h)->add_receiver(r);
  });
}

int insem_model_clear_receivers(void* h) {
  return guard([&] { ref<Model>(h)->clear_receivers(); });
}

int insem_model_set_time(void* h, double dt, double safety, int32_t n_steps, double duration, int32_t record_every,
                         int32_t energy_every) {
  return guard([&] {
    TimeConfig t;
    t.dt = dt;
    t.safety = safety;
    t.n_steps = n_steps;
    t.duration = duration;
    t.record_every = record_every;
    t.energy_every = energy_every;
    ref<Model>(h)->set_time(t);
  });
}

int insem_model_setup(void* h) {
  return guard([&] { ref<Model>(h)->setup(); });
}

int insem_model_info(void* h, int32_t* out) {
  return guard([&] {
    const Model& m = *ref<Model>(h);
    out[0] = m.forest().n_leaves();
    out[1] = m.forest().max_level();
    const auto c = m.kind_counts();
    out[2] = c.at("rect");
    out[3] = c.at("affine");
    out[4] = c.at("curved");
  });
}

int insem_model_leaf_points(void* h, double* x, double* z) {
  return guard([&] {
    const Geometry& g = ref<Model>(h)->geometry();
    std::memcpy(x, g.x.data(), g.x.size() * sizeof(double));
    std::memcpy(z, g.z.data(), g.z.size() * sizeof(double));
  });
}

void* insem_model_build(void* h) {
  return guard_ptr([&]() -> void* { return ref<Model>(h)->build().release(); });
}

// ----------------------------------------------------------------- image

void insem_image_free(void* h) { delete static_cast<SetupImage*>(h); }

const insem_hip_setup* insem_image_view(void* h) {
  return static_cast<const insem_hip_setup*>(guard_ptr([&]() -> void* { return &ref<SetupImage>(h)->view; }));
}

int insem_image_domain_info(void* h, int32_t di, int32_t* out) {
  return guard([&] {
    const SetupImage& im = *ref<SetupImage>(h);
    INSEM_REQUIRE(di >= 0 && di < static_cast<int>(im.domains.size()), "invalid domain " << di);
    const auto& d = im.domains[di];
    out[0] = static_cast<int32_t>(d.medium);
    out[1] = d.ncomp;
    out[2] = d.dofs.n_global;
    out[3] = d.dofs.n_elements;
    out[4] = static_cast<int32_t>(d.blocks.size());
    out[5] = static_cast<int32_t>(d.dirichlet_points.size());
    out[6] = static_cast<int32_t>(d.absorbing_points.size());
    out[7] = static_cast<int32_t>(d.axis_points.size());
    out[8] = d.dofs.has_constraints() ? 1 : 0;
    out[9] = d.azimuthal_order;
  });
}

int insem_image_block_info(void* h, int32_t di, int32_t bi, int32_t* out) {
  return guard([&] {
    const auto& d = ref<SetupImage>(h)->domains.at(di);
    const auto& b = d.blocks.at(bi);
    out[0] = b.kind;
    out[1] = static_cast<int32_t>(b.elements.size());
    out[2] = b.azimuthal_order;
  });
}

int insem_image_n_domains(void* h) {
  int n = -1;
  (void)guard([&] { n = static_cast<int>(ref<SetupImage>(h)->domains.size()); });
  return n;
}

int insem_image_coords(void* h, int32_t di, double* out) {
  return guard([&] {
    const auto& d = ref<SetupImage>(h)->domains.at(di);
    std::memcpy(out, d.coords.data(), d.coords.size() * sizeof(double));
  });
}

int insem_image_mass(void* h, int32_t di, double* out) {
  return guard([&] {
    const auto& d = ref<SetupImage>(h)->domains.at(di);
    for (std::size_t g = 0; g < d.mass.size(); ++g) out[g] = d.mass[g];
  });
}

int64_t insem_image_start_len(void* h) {
  int64_t n = -1;
  (void)guard([&] {
    const SetupImage& im = *ref<SetupImage>(h);
    n = 0;
    for (const auto& d : im.domains) n += static_cast<int64_t>(d.ncomp) * d.dofs.n_global;
    if (!im.coupling.fluid_points.empty()) n *= 2;
  });
  return n;
}

int insem_image_start_vectors(void* h, float* out) {
  return guard([&] {
    const auto v = ref<SetupImage>(h)->power_start_vectors();
    std::memcpy(out, v.data(), v.size() * sizeof(float));
  });
}

int insem_image_amplitudes(void* h, double dt, int32_t n_steps, float* out) {
  return guard([&] {
    const auto a = ref<SetupImage>(h)->source_amplitudes(dt, n_steps);
    std::memcpy(out, a.data(), a.size() * sizeof(float));
  });
}

int insem_image_time(void* h, double* d, int32_t* i) {
  return guard([&] {
    const TimeConfig& t = ref<SetupImage>(h)->time;
    d[0] = t.dt;
    d[1] = t.safety;
    d[2] = t.duration;
    i[0] = t.n_steps;
    i[1] = t.record_every;
    i[2] = t.energy_every;
  });
}

// The time step and step count a model gets from lambda_max (as its Solver
// would): d = dt, stable_dt; *n = n_steps.
int insem_image_model_time(void* h, double lambda, double* d, int32_t* n) {
  return guard([&] {
    const ModelTime m = model_time(ref<SetupImage>(h)->time, lambda);
    d[0] = m.dt;
    d[1] = m.stable_dt;
    *n = m.n_steps;
  });
}

int insem_image_n_traces(void* h) {
  int n = -1;
  (void)guard([&] {
    n = 0;
    for (const auto& r : ref<SetupImage>(h)->receivers) n += static_cast<int>(r.quantities.size());
  });
  return n;
}

int insem_image_trace_info(void* h, int32_t t, char* name, int32_t name_len, char* quantity, int32_t q_len,
                           int32_t* ncomp) {
  return guard([&] {
    int k = 0;
    for (const auto& r : ref<SetupImage>(h)->receivers)
      for (size_t q = 0; q < r.quantities.size(); ++q, ++k)
        if (k == t) {
          std::snprintf(name, name_len, "%s", r.name.c_str());
          std::snprintf(quantity, q_len, "%s", r.quantities[q].c_str());
          *ncomp = r.ncomp[q];
          return;
        }
    throw Error("trace index out of range");
  });
}

int insem_image_counts(void* h, int64_t* out) {
  return guard([&] {
    const SetupImage& im = *ref<SetupImage>(h);
    out[0] = im.total_points();
    out[1] = im.total_dofs();
    out[2] = static_cast<int64_t>(im.coupling.fluid_points.size());
    out[3] = static_cast<int64_t>(im.sources.size());
    out[4] = static_cast<int64_t>(im.receivers.size());
    out[5] = im.ngll;
    out[6] = im.axisym ? 1 : 0;
    out[7] = im.azimuthal_order;
  });
}

// ---------------------------------------------------------------- solver

void* insem_solver_new(void* image) {
  return guard_ptr([&]() -> void* { return new Solver(std::unique_ptr<SetupImage>(static_cast<SetupImage*>(image))); });
}

void insem_solver_free(void* h) { delete static_cast<Solver*>(h); }

void* insem_solver_image(void* h) {
  return guard_ptr([&]() -> void* { return const_cast<SetupImage*>(&ref<Solver>(h)->image()); });
}

void* insem_solver_context(void* h) {
  return guard_ptr([&]() -> void* { return ref<Solver>(h)->context(); });
}

int insem_solver_info(void* h, double* d, int32_t* i) {
  return guard([&] {
    const Solver& s = *ref<Solver>(h);
    d[0] = s.dt();
    d[1] = s.stable_dt();
    i[0] = s.n_steps();
    i[1] = s.step();
  });
}

int insem_solver_advance(void* h, int32_t n) {
  return guard([&] { ref<Solver>(h)->advance(n); });
}

int insem_solver_set_state(void* h, int32_t medium, const double* u, const double* v, int64_t n) {
  return guard([&] {
    ref<Solver>(h)->set_state(medium == 0 ? Medium::Acoustic : Medium::Elastic, std::vector<double>(u, u + n),
                                       std::vector<double>(v, v + n));
  });
}

int insem_solver_field(void* h, int32_t di, int32_t which, double* out) {
  return guard([&] {
    const auto f = ref<Solver>(h)->field(di, which);
    std::memcpy(out, f.data(), f.size() * sizeof(double));
  });
}

int insem_solver_n_samples(void* h) {
  int32_t n = -1;
  (void)guard([&] { insem_hip_n_samples(ref<Solver>(h)->context(), &n); });
  return n;
}

int insem_solver_traces(void* h, double* out) {
  return guard([&] {
    std::size_t o = 0;
    for (const auto& t : ref<Solver>(h)->traces()) {
      std::memcpy(out + o, t.data.data(), t.data.size() * sizeof(double));
      o += t.data.size();
    }
  });
}

int insem_solver_n_energy(void* h) {
  int32_t n = -1;
  (void)guard([&] { insem_hip_n_energy(ref<Solver>(h)->context(), &n); });
  return n;
}

int insem_solver_energy(void* h, double* out) {
  return guard([&] {
    const auto e = ref<Solver>(h)->energy();
    for (std::size_t i = 0; i < e.size(); ++i)
      for (int k = 0; k < 3; ++k) out[3 * i + k] = e[i][k];
  });
}

// ---------------------------------------------------------------- batch

// Takes ownership of the n images.
void* insem_batch_new(void* const* images, int32_t n) {
  return guard_ptr([&]() -> void* {
    std::vector<std::unique_ptr<SetupImage>> v;
    for (int32_t i = 0; i < n; ++i) v.emplace_back(static_cast<SetupImage*>(images[i]));
    return new BatchSolver(std::move(v));
  });
}

// As insem_batch_new, with each model's lambda_max given (from a power job
// of it alone): makes no power iteration.
void* insem_batch_new_lambdas(void* const* images, const double* lambdas, int32_t n) {
  return guard_ptr([&]() -> void* {
    INSEM_REQUIRE(lambdas && n >= 0, "null lambda_max values");
    std::vector<std::unique_ptr<SetupImage>> v;
    for (int32_t i = 0; i < n; ++i) v.emplace_back(static_cast<SetupImage*>(images[i]));
    return new BatchSolver(std::move(v), std::vector<double>(lambdas, lambdas + n));
  });
}

void insem_batch_free(void* h) { delete static_cast<BatchSolver*>(h); }

// ---------------------------------------------------------------- power jobs

// Uploads the model to a context of its own and queues its power
// iteration; the image stays the caller's and may be freed afterwards.
void* insem_power_start(void* image) {
  return guard_ptr([&]() -> void* { return new PowerJob(*static_cast<const SetupImage*>(image)); });
}

// Device bytes of the model's context (host only, any thread).
int insem_image_stage_size(void* image, int64_t* bytes) {
  return guard([&] { check_hip(insem_hip_stage_size(&ref<SetupImage>(image)->view, bytes)); });
}

// The host half of a power iteration's start for a block from
// insem_hip_reserve (host only, any thread); null on failure (the block
// then stays reserved).
void* insem_power_stage(void* image, void* block) {
  return guard_ptr([&]() -> void* { return new StagedPower(*ref<SetupImage>(image), block); });
}

// Device half: realizes the staged context and starts the iteration.
void* insem_power_start_staged(void* staged) {
  return guard_ptr([&]() -> void* { return new PowerJob(*ref<StagedPower>(staged)); });
}

// Frees a staged start (a context never realized gives its block back).
void insem_power_staged_free(void* staged) { delete static_cast<StagedPower*>(staged); }

// *ready = 1 when the iteration has finished (lambda then returns at once).
int insem_power_ready(void* job, int32_t* ready) {
  return guard([&] { *ready = static_cast<PowerJob*>(job)->ready() ? 1 : 0; });
}

// Waits for lambda_max; the job's context is freed, the job stays readable.
int insem_power_lambda(void* job, double* lambda) {
  return guard([&] { *lambda = static_cast<PowerJob*>(job)->lambda(); });
}

void insem_power_free(void* job) { delete static_cast<PowerJob*>(job); }

// A Solver of the model on the job's own context (the job must have been
// started from this image, which the solver takes over).
void* insem_power_solver(void* job, void* image) {
  return guard_ptr([&]() -> void* {
    std::unique_ptr<SetupImage> img(static_cast<SetupImage*>(image));
    double lambda = 0.0;
    insem_hip_ctx* ctx = static_cast<PowerJob*>(job)->take_context(&lambda);
    try {
      return new Solver(std::move(img), ctx, lambda);
    } catch (...) {
      insem_hip_destroy(ctx);
      throw;
    }
  });
}

void* insem_batch_context(void* h) {
  return guard_ptr([&]() -> void* { return ref<BatchSolver>(h)->context(); });
}

void* insem_batch_model_image(void* h, int32_t m) {
  return guard_ptr([&]() -> void* { return const_cast<SetupImage*>(&ref<BatchSolver>(h)->model(m)); });
}

// d: dt, stable_dt; i: n_steps, samples so far.  m = -1: n_models, max steps, union step.
int insem_batch_info(void* h, int32_t m, double* d, int32_t* i) {
  return guard([&] {
    const BatchSolver& b = *ref<BatchSolver>(h);
    if (m < 0) {
      i[0] = b.n_models();
      i[1] = b.max_steps();
      i[2] = b.step();
      return;
    }
    INSEM_REQUIRE(m < b.n_models(), "no model " << m << " in the batch");
    d[0] = b.dt(m);
    d[1] = b.stable_dt(m);
    i[0] = b.n_steps(m);
    i[1] = b.n_samples(m);
  });
}

int insem_batch_advance(void* h, int32_t n) {
  return guard([&] { ref<BatchSolver>(h)->advance(n); });
}

int insem_batch_traces(void* h, int32_t m, double* out) {
  return guard([&] {
    const BatchSolver& b = *ref<BatchSolver>(h);
    INSEM_REQUIRE(m >= 0 && m < b.n_models(), "no model " << m << " in the batch");
    std::size_t o = 0;
    for (const auto& t : b.traces(m)) {
      std::memcpy(out + o, t.data.data(), t.data.size() * sizeof(double));
      o += t.data.size();
    }
  });
}

int insem_batch_traces_done(void* h, int32_t m, double* out) {
  return guard([&] {
    const BatchSolver& b = *ref<BatchSolver>(h);
    INSEM_REQUIRE(m >= 0 && m < b.n_models(), "no model " << m << " in the batch");
    std::size_t o = 0;
    for (const auto& t : b.traces_done(m)) {
      std::memcpy(out + o, t.data.data(), t.data.size() * sizeof(double));
      o += t.data.size();
    }
  });
}

int insem_batch_device_step(void* h, int32_t* step) {
  return guard([&] { *step = ref<BatchSolver>(h)->device_step(); });
}

int64_t insem_batch_footprint(void* h) {
  try {
    return ref<BatchSolver>(h)->footprint();
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
