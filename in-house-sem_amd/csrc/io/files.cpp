#include "files.hpp"

#include <algorithm>
#include <set>

#include "../host/error.hpp"
#include "h5file.hpp"

namespace insem_amd {
namespace {

constexpr const char* kCaseFormat = "insem-case";
constexpr const char* kResultsFormat = "insem-results";
constexpr int64_t kFormatVersion = 1;

std::string joined(const std::vector<std::string>& v) {
  std::string s;
  for (std::size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + v[i];
  return s;
}

std::vector<std::string> items(const std::string& s) {
  std::vector<std::string> out;
  std::size_t a = 0;
  while (a <= s.size()) {
    const std::size_t b = std::min(s.find(',', a), s.size());
    if (b > a) out.push_back(s.substr(a, b - a));
    a = b + 1;
  }
  return out;
}

std::string num(std::size_t i) { return std::to_string(i); }

Stf stf_from(const h5::File& f, const std::string& g, std::size_t index) {
  const std::string kind = f.get_string(g, "stf");
  const double amp = f.get_double(g, "amplitude");
  if (kind == "samples") {
    Stf s;
    s.type = Stf::Samples;
    s.samples = f.get<double>(g + "/stf_samples");
    s.sample_dt = f.get_double(g, "sample_dt");
    s.amplitude = amp;
    INSEM_REQUIRE(s.sample_dt > 0.0 && !s.samples.empty(), "source " << index << " has no usable samples");
    return s;
  }
  for (int t : {Stf::Ricker, Stf::Gaussian, Stf::GaussianDerivative})
    if (kind == stf_name(t)) return Stf::make(t, f.get_double(g, "f0"), f.get_double(g, "t0"), amp);
  throw Error("source " + num(index) + " of '" + f.path() + "' has the unknown time function '" + kind + "'");
}

}  // namespace

void save_case(const Model& model, const std::string& path) {
  h5::File f(path, h5::Open::Create);
  f.set("/", "format", kCaseFormat);
  f.set("/", "version", kFormatVersion);
  f.set("/", "insem_version", insem_version());

  const MacroMesh& m = model.mesh();
  const std::size_t nn = static_cast<std::size_t>(m.n_nodes()), nc = static_cast<std::size_t>(m.n_cells());
  f.put("/mesh/nodes", m.nodes().data(), {nn, 2});
  f.put("/mesh/cells", m.cells().data(), {nc, static_cast<std::size_t>(m.nodes_per_cell())});
  const std::vector<int32_t> material(m.materials().begin(), m.materials().end());
  f.put("/mesh/material", material.data(), {nc});
  std::vector<int32_t> faces;
  for (const BoundaryFace& b : m.boundary()) faces.insert(faces.end(), {b.cell, b.face, b.tag});
  f.put("/mesh/boundary", faces.data(), {m.boundary().size(), 3});
  f.set("/mesh", "cell_type", m.nodes_per_cell() == 4 ? "quad4" : "quad9");
  f.set("/mesh", "coords", m.axisymmetric() ? "axisymmetric" : "cartesian");

  const Discretization& d = model.discretization();
  f.make_group("/discretization");
  f.set("/discretization", "order", static_cast<int64_t>(d.order));
  f.set("/discretization", "uniform_level", static_cast<int64_t>(d.uniform_level));
  f.set("/discretization", "f_max", d.f_max);
  f.set("/discretization", "points_per_wavelength", d.points_per_wavelength);
  f.set("/discretization", "max_level", static_cast<int64_t>(d.max_level));
  f.set("/discretization", "azimuthal_order", static_cast<int64_t>(d.azimuthal_order));
  std::vector<double> regions;
  for (const Region& r : d.regions) regions.insert(regions.end(), {r.xmin, r.xmax, r.zmin, r.zmax, double(r.level)});
  f.put("/discretization/regions", regions.data(), {d.regions.size(), 5});

  f.make_group("/materials");
  for (const auto& [region, mat] : model.materials()) {
    const std::string g = "/materials/" + std::to_string(region);
    f.set(g, "medium", medium_name(mat.medium));
    f.set(g, "vp", mat.vp);
    f.set(g, "vs", mat.vs);
    f.set(g, "rho", mat.rho);
  }

  f.make_group("/boundary_conditions");
  for (const auto& [tag, type] : model.boundaries()) f.set("/boundary_conditions", std::to_string(tag), boundary_name(type));

  const auto& sources = model.sources();
  f.make_group("/sources");
  f.set("/sources", "count", static_cast<int64_t>(sources.size()));
  for (std::size_t i = 0; i < sources.size(); ++i) {
    const Source& s = sources[i];
    const std::string g = "/sources/" + num(i);
    f.put(g + "/position", std::vector<double>{s.position[0], s.position[1]});
    f.put(g + "/components", std::vector<double>(s.components.begin(), s.components.end()));
    f.set(g, "kind", s.type_name());
    f.set(g, "stf", stf_name(s.stf.type));
    f.set(g, "f0", s.stf.f0);
    f.set(g, "t0", s.stf.t0);
    f.set(g, "amplitude", s.stf.amplitude);
    if (s.stf.type == Stf::Samples) {
      f.put(g + "/stf_samples", s.stf.samples);
      f.set(g, "sample_dt", s.stf.sample_dt);
    }
  }

  const auto& receivers = model.receivers();
  f.make_group("/receivers");
  f.set("/receivers", "count", static_cast<int64_t>(receivers.size()));
  for (std::size_t i = 0; i < receivers.size(); ++i) {
    const Receiver& r = receivers[i];
    const std::string g = "/receivers/" + num(i);
    f.put(g + "/position", std::vector<double>{r.position[0], r.position[1]});
    f.set(g, "name", r.name);
    f.set(g, "quantities", joined(r.quantities));
    f.set(g, "medium", r.medium);
  }

  const TimeConfig& t = model.time();
  f.make_group("/time");
  f.set("/time", "dt", t.dt);
  f.set("/time", "safety", t.safety);
  f.set("/time", "n_steps", static_cast<int64_t>(t.n_steps));
  f.set("/time", "duration", t.duration);
  f.set("/time", "record_every", static_cast<int64_t>(t.record_every));
  f.set("/time", "energy_every", static_cast<int64_t>(t.energy_every));
  f.set("/time", "snapshot_every", static_cast<int64_t>(t.snapshot_every));
  f.set("/time", "snapshot_quantities", joined(t.snapshot_quantities));
}

std::unique_ptr<Model> load_case(const std::string& path) {
  const h5::File f(path, h5::Open::Read);
  INSEM_REQUIRE(f.has_attr("/", "format") && f.get_string("/", "format") == kCaseFormat,
                "'" << path << "' is not an insem case file");
  const int64_t version = f.get_int("/", "version");
  INSEM_REQUIRE(version == kFormatVersion,
                "'" << path << "' is a case file of version " << version << "; this reader takes version " << kFormatVersion);

  std::vector<std::size_t> shape;
  std::vector<double> nodes = f.get<double>("/mesh/nodes", &shape);
  INSEM_REQUIRE(shape.size() == 2 && shape[1] == 2, "/mesh/nodes of '" << path << "' is not an (n, 2) array");
  std::vector<int32_t> cells = f.get<int32_t>("/mesh/cells");
  const std::vector<int32_t> material = f.get<int32_t>("/mesh/material");
  const std::vector<int32_t> faces = f.get<int32_t>("/mesh/boundary");
  INSEM_REQUIRE(faces.size() % 3 == 0, "/mesh/boundary of '" << path << "' is not an (n, 3) array");
  std::vector<BoundaryFace> boundary;
  for (std::size_t i = 0; i < faces.size(); i += 3) boundary.push_back({faces[i], faces[i + 1], faces[i + 2]});
  const std::string cell_type = f.get_string("/mesh", "cell_type");
  const std::string coords = f.get_string("/mesh", "coords");
  INSEM_REQUIRE(cell_type == "quad4" || cell_type == "quad9", "unknown cell type '" << cell_type << "'");
  INSEM_REQUIRE(coords == "cartesian" || coords == "axisymmetric", "unknown coordinate system '" << coords << "'");
  auto mesh = std::make_shared<const MacroMesh>(cell_type == "quad4" ? 4 : 9, coords == "axisymmetric", std::move(nodes),
                                                std::move(cells), std::vector<int>(material.begin(), material.end()),
                                                std::move(boundary));

  Discretization disc;
  disc.order = static_cast<int>(f.get_int("/discretization", "order"));
  disc.uniform_level = static_cast<int>(f.get_int("/discretization", "uniform_level"));
  if (f.has_attr("/discretization", "f_max")) {
    disc.f_max = f.get_double("/discretization", "f_max");
    disc.points_per_wavelength = f.get_double("/discretization", "points_per_wavelength");
    disc.max_level = static_cast<int>(f.get_int("/discretization", "max_level"));
    const std::vector<double> r = f.get<double>("/discretization/regions");
    for (std::size_t i = 0; i + 5 <= r.size(); i += 5)
      disc.regions.push_back({r[i], r[i + 1], r[i + 2], r[i + 3], static_cast<int>(r[i + 4])});
  }
  if (f.has_attr("/discretization", "azimuthal_order"))
    disc.azimuthal_order = static_cast<int>(f.get_int("/discretization", "azimuthal_order"));
  auto model = std::make_unique<Model>(mesh, disc);

  if (f.has("/materials"))
    for (int region : std::set<int>(mesh->materials().begin(), mesh->materials().end())) {
      const std::string g = "/materials/" + std::to_string(region);
      if (!f.has(g)) continue;
      Material mat;
      mat.medium = medium_from_name(f.get_string(g, "medium"));
      mat.vp = f.get_double(g, "vp");
      mat.vs = f.get_double(g, "vs");
      mat.rho = f.get_double(g, "rho");
      model->set_material(region, mat);
    }
  if (f.has("/boundary_conditions")) {
    std::set<int> tags;
    for (const BoundaryFace& b : mesh->boundary()) tags.insert(b.tag);
    for (int tag : tags)
      if (f.has_attr("/boundary_conditions", std::to_string(tag)))
        model->set_boundary(tag, boundary_from_name(f.get_string("/boundary_conditions", std::to_string(tag))));
  }
  if (f.has("/sources")) {
    const int64_t n = f.get_int("/sources", "count");
    for (int64_t i = 0; i < n; ++i) {
      const std::string g = "/sources/" + num(static_cast<std::size_t>(i));
      const std::vector<double> pos = f.get<double>(g + "/position"), comp = f.get<double>(g + "/components");
      INSEM_REQUIRE(pos.size() == 2 && comp.size() == 3, "source " << i << " of '" << path << "' is malformed");
      Source s;
      const std::string kind = f.get_string(g, "kind");
      s.position = {pos[0], pos[1]};
      if (kind == "force" || kind == "dipole") {
        s.type = kind == "force" ? Source::Force : Source::Dipole;
        s.components = {comp[0], comp[1], 0.0};
      } else if (kind == "moment_tensor") {
        s.type = Source::MomentTensor;
        s.components = {comp[0], comp[1], comp[2]};
      } else if (kind == "monopole") {
        s.type = Source::Monopole;
        s.components = {1.0, 0.0, 0.0};
      } else {
        throw Error("source " + std::to_string(i) + " of '" + path + "' is of the unknown kind '" + kind + "'");
      }
      s.stf = stf_from(f, g, static_cast<std::size_t>(i));
      model->add_source(s);
    }
  }
  if (f.has("/receivers")) {
    const int64_t n = f.get_int("/receivers", "count");
    for (int64_t i = 0; i < n; ++i) {
      const std::string g = "/receivers/" + num(static_cast<std::size_t>(i));
      const std::vector<double> pos = f.get<double>(g + "/position");
      INSEM_REQUIRE(pos.size() == 2, "receiver " << i << " of '" << path << "' is malformed");
      Receiver r;
      r.name = f.get_string(g, "name");
      r.position = {pos[0], pos[1]};
      r.quantities = items(f.get_string(g, "quantities"));
      r.medium = f.has_attr(g, "medium") ? f.get_string(g, "medium") : "";
      model->add_receiver(r);
    }
  }
  if (f.has("/time")) {
    TimeConfig t;
    t.dt = f.get_double("/time", "dt");
    t.safety = f.get_double("/time", "safety");
    t.n_steps = static_cast<int>(f.get_int("/time", "n_steps"));
    t.duration = f.get_double("/time", "duration");
    t.record_every = static_cast<int>(f.get_int("/time", "record_every"));
    t.energy_every = static_cast<int>(f.get_int("/time", "energy_every"));
    if (f.has_attr("/time", "snapshot_every")) {
      t.snapshot_every = static_cast<int>(f.get_int("/time", "snapshot_every"));
      t.snapshot_quantities = items(f.get_string("/time", "snapshot_quantities"));
    }
    model->set_time(t);
  }
  return model;
}

void write_results(const Solver& s, const std::string& path) {
  const SetupImage& img = s.image();
  h5::File f(path, h5::Open::Create);
  f.set("/", "format", kResultsFormat);
  f.set("/", "version", kFormatVersion);
  f.set("/", "insem_version", insem_version());
  f.set("/", "dt", s.dt());
  f.set("/", "n_steps", static_cast<int64_t>(s.n_steps()));
  f.set("/", "steps_done", static_cast<int64_t>(s.step()));
  f.set("/", "record_every", static_cast<int64_t>(img.time.record_every));
  f.set("/", "coords", img.axisym ? "axisymmetric" : "cartesian");
  f.set("/", "azimuthal_order", static_cast<int64_t>(img.azimuthal_order));
  f.put("/time", s.sample_times());
  for (const Solver::Trace& t : s.traces()) {
    const std::size_t n = t.data.size() / static_cast<std::size_t>(t.ncomp);
    f.put("/receivers/" + t.receiver + "/" + t.quantity, t.data.data(), {n, static_cast<std::size_t>(t.ncomp)});
  }
  for (const auto& r : img.receivers) {
    const std::string g = "/receivers/" + r.name;
    f.set(g, "medium", medium_name(img.domains[r.domain].medium));
    f.set(g, "x", r.position[0]);
    f.set(g, "z", r.position[1]);
  }
  if (!s.snapshots().empty()) {
    const std::vector<double>& times = s.snapshot_times();
    f.put("/snapshots/time", times);
    for (Medium m : {Medium::Acoustic, Medium::Elastic}) {
      if (img.domain_index(m) < 0) continue;
      const std::string g = std::string("/snapshots/") + medium_name(m);
      const std::vector<double> p = s.snapshot_points(m);
      const std::size_t n = p.size() / 2;
      std::vector<double> xz(2 * n);
      for (std::size_t i = 0; i < n; ++i) {
        xz[2 * i] = p[i];
        xz[2 * i + 1] = p[n + i];
      }
      f.put(g + "/points", xz.data(), {n, 2});
      const std::vector<int32_t> cells = s.snapshot_cells(m);
      f.put(g + "/cells", cells.data(), {cells.size() / 4, 4});
    }
    for (const Solver::Snapshot& sn : s.snapshots()) {
      const std::size_t ns = times.size();
      const std::size_t n = ns ? sn.data.size() / (ns * static_cast<std::size_t>(sn.ncomp)) : 0;
      f.put(std::string("/snapshots/") + medium_name(sn.medium) + "/" + sn.quantity, sn.data.data(),
            {ns, static_cast<std::size_t>(sn.ncomp), n});
    }
  }
  const auto energy = s.energy();
  if (!energy.empty()) {
    std::vector<double> e;
    for (const auto& row : energy) e.insert(e.end(), row.begin(), row.end());
    f.put("/energy", e.data(), {energy.size(), 3});
    f.set("/energy", "columns", "time,kinetic,potential");
  }
}

}  // namespace insem_amd
