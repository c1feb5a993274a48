// insem_amd._core: the reference's compiled Python module (insem._core,
// python/src/bindings.cpp:110-494) on this backend.  It exports every name
// the reference's Python package imports from its core (python/insem/
// __init__.py:7-29, basis.py:3-10, config.py, cli.py, mesh/templates.py),
// with the same signatures, argument names, array shapes and errors, over
// the host library (mesh, forest, point maps, geometry, model description,
// case and results files) and the GPU solver behind the C ABI.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../host/basis.hpp"
#include "../host/dofmap.hpp"
#include "../host/error.hpp"
#include "../host/geometry.hpp"
#include "../host/mesh.hpp"
#include "../host/model.hpp"
#include "../host/solver.hpp"
#include "../io/files.hpp"
#include "../io/h5file.hpp"

namespace py = pybind11;
using namespace insem_amd;

namespace {

template <class T>
using In = py::array_t<T, py::array::c_style | py::array::forcecast>;

template <class T>
py::array_t<T> array(const std::vector<T>& v, std::vector<py::ssize_t> shape) {
  py::array_t<T> a(shape);
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

template <class T>
py::array_t<T> array(const std::vector<T>& v) {
  return array(v, {static_cast<py::ssize_t>(v.size())});
}

template <class T>
std::vector<T> values(const In<T>& a) {
  return std::vector<T>(a.data(), a.data() + a.size());
}

py::tuple rule_tuple(const Rule& r) { return py::make_tuple(array(r.points), array(r.weights)); }

// P_n^(a, b)(x) from the explicit sum
//   sum_s C(n + a, n - s) C(n + b, s) ((x - 1) / 2)^s ((x + 1) / 2)^(n - s)
// (Szego, Orthogonal Polynomials, eq. 4.3.2), generalised binomials.
double jacobi_value(int n, double a, double b, double x) {
  INSEM_REQUIRE(n >= 0, "the degree of a Jacobi polynomial is at least 0, not " << n);
  auto binom = [](double top, int k) {
    double c = 1.0;
    for (int i = 1; i <= k; ++i) c *= (top - k + i) / i;
    return c;
  };
  const double lo = 0.5 * (x - 1.0), hi = 0.5 * (x + 1.0);
  double sum = 0.0;
  for (int s = 0; s <= n; ++s) sum += binom(n + a, n - s) * binom(n + b, s) * std::pow(lo, s) * std::pow(hi, n - s);
  return sum;
}

Medium medium_arg(const std::string& s) { return medium_from_name(s); }

std::shared_ptr<MacroMesh> new_mesh(const std::string& cell_type, const std::string& coords, const In<double>& nodes,
                                    const In<int32_t>& cells, const In<int32_t>& material, const In<int32_t>& boundary) {
  INSEM_REQUIRE(cell_type == "quad4" || cell_type == "quad9", "unknown cell type '" << cell_type << "'");
  INSEM_REQUIRE(coords == "cartesian" || coords == "axisymmetric", "unknown coordinate system '" << coords << "'");
  const int npc = cell_type == "quad4" ? 4 : 9;
  INSEM_REQUIRE(nodes.ndim() == 2 && nodes.shape(1) == 2, "nodes must have shape (n_nodes, 2)");
  INSEM_REQUIRE(cells.ndim() == 2 && cells.shape(1) == npc, "cells must have shape (n_cells, " << npc << ") for " << cell_type);
  INSEM_REQUIRE(material.ndim() == 1 && material.shape(0) == cells.shape(0), "material must have shape (n_cells,)");
  INSEM_REQUIRE(boundary.size() == 0 || (boundary.ndim() == 2 && boundary.shape(1) == 3),
                "boundary must have shape (n_faces, 3): (cell, face, tag)");
  std::vector<BoundaryFace> faces;
  const int32_t* b = boundary.data();
  for (py::ssize_t i = 0; i < boundary.size() / 3; ++i) faces.push_back({b[3 * i], b[3 * i + 1], b[3 * i + 2]});
  const std::vector<int32_t> m = values(material);
  return std::make_shared<MacroMesh>(npc, coords == "axisymmetric", values(nodes), values(cells),
                                     std::vector<int>(m.begin(), m.end()), std::move(faces));
}

// (n_global,) or (ncomp, n_global) -> (n_elements, [ncomp,] ngll, ngll):
// every element point takes its global value, or its constraint row's
// weighted sum (in double).
py::array_t<double> dof_gather(const DofMap& d, const In<double>& g) {
  const bool scalar = g.ndim() == 1;
  INSEM_REQUIRE((scalar && g.shape(0) == d.n_global) || (g.ndim() == 2 && g.shape(1) == d.n_global),
                "global field must have shape (n_global,) or (ncomp, n_global)");
  const int nc = scalar ? 1 : static_cast<int>(g.shape(0));
  std::vector<py::ssize_t> shape{d.n_elements};
  if (!scalar) shape.push_back(nc);
  shape.insert(shape.end(), {d.ngll, d.ngll});
  py::array_t<double> out(shape);
  double* o = out.mutable_data();
  const double* in = g.data();
  const std::size_t ng = static_cast<std::size_t>(d.n_global), npe = static_cast<std::size_t>(d.npe);
  for (int32_t e = 0; e < d.n_elements; ++e)
    for (int c = 0; c < nc; ++c)
      for (std::size_t k = 0; k < npe; ++k) {
        const int32_t m = d.l2g[e * npe + k];
        double v;
        if (m >= 0) {
          v = in[c * ng + m];
        } else {
          v = 0.0;
          for (int32_t q = d.row_offsets[-1 - m]; q < d.row_offsets[-m]; ++q) v += d.row_weights[q] * in[c * ng + d.row_dofs[q]];
        }
        o[(e * nc + c) * npe + k] = v;
      }
  return out;
}

// The transpose: each global point sums its contributing slots (weighted
// where constrained) in ascending slot order.
py::array_t<double> dof_assemble(const DofMap& d, const In<double>& l) {
  const bool scalar = l.ndim() == 3;
  INSEM_REQUIRE((scalar || l.ndim() == 4) && l.shape(0) == d.n_elements && l.shape(l.ndim() - 1) == d.ngll &&
                    l.shape(l.ndim() - 2) == d.ngll,
                "local field must have shape (n_elements, [ncomp,] ngll, ngll)");
  const int nc = scalar ? 1 : static_cast<int>(l.shape(1));
  std::vector<py::ssize_t> shape;
  if (!scalar) shape.push_back(nc);
  shape.push_back(d.n_global);
  py::array_t<double> out(shape);
  double* o = out.mutable_data();
  const double* in = l.data();
  const std::size_t ng = static_cast<std::size_t>(d.n_global), npe = static_cast<std::size_t>(d.npe);
  const bool weighted = !d.t_weights.empty();
  for (int c = 0; c < nc; ++c)
    for (int32_t g = 0; g < d.n_global; ++g) {
      double s = 0.0;
      for (int32_t q = d.t_offsets[g]; q < d.t_offsets[g + 1]; ++q) {
        const std::size_t slot = static_cast<std::size_t>(d.t_slots[q]);
        const double v = in[((slot / npe) * nc + c) * npe + slot % npe];
        s += weighted ? d.t_weights[q] * v : v;
      }
      o[c * ng + g] = s;
    }
  return out;
}

// The reference's Simulation: a model description, its set-up, and the
// solver built from it on first use (any change of the description drops
// the solver, and a change that moves the refinement drops the set-up).
struct Sim {
  std::unique_ptr<Model> model;
  std::unique_ptr<Solver> solver;
  Solver& get_solver() {
    if (!solver) {
      model->setup();
      solver = std::make_unique<Solver>(model->build());
    }
    return *solver;
  }
};

std::string material_repr(const Material& m) {
  std::ostringstream os;
  os << "<Material " << medium_name(m.medium) << " vp=" << m.vp;
  if (m.medium == Medium::Elastic) os << " vs=" << m.vs;
  os << " rho=" << m.rho << ">";
  return os.str();
}

using RegionTuple = std::tuple<double, double, double, double, int>;

const SetupImage::Domain& solver_domain(const Solver& s, const std::string& medium) {
  const int di = s.domain_index(medium_arg(medium));
  INSEM_REQUIRE(di >= 0, "the model has no " << medium << " region");
  return s.image().domains[di];
}

}  // namespace

PYBIND11_MODULE(_core, m) {
  m.doc() = "insem core on the MI355X backend";
  m.attr("__version__") = insem_version();
  m.attr("real_dtype") = "float32";
  py::register_exception<Error>(m, "InsemError", PyExc_RuntimeError);

  // ---- basis
  m.def("gll", [](int n) { return rule_tuple(gll_rule(n)); }, py::arg("n_points"),
        "Gauss-Lobatto-Legendre points and weights on [-1, 1].");
  m.def("glj01", [](int n) { return rule_tuple(glj01_rule(n)); }, py::arg("n_points"),
        "Gauss-Lobatto-Jacobi (0,1) points and weights (weight 1 + x) on [-1, 1].");
  m.def("gauss_lobatto_jacobi", [](int n, int a, int b) { return rule_tuple(lobatto_rule(n, a, b)); },
        py::arg("n_points"), py::arg("alpha"), py::arg("beta"));
  m.def("jacobi", &jacobi_value, py::arg("n"), py::arg("a"), py::arg("b"), py::arg("x"));
  m.def(
      "derivative_matrix",
      [](const In<double>& nodes) {
        const Lagrange1D L(values(nodes));
        return array(L.derivative_matrix(), {L.size(), L.size()});
      },
      py::arg("nodes"), "D[i, j] = l_j'(x_i) for the Lagrange basis on `nodes`.");
  m.def(
      "interpolation_matrix",
      [](const In<double>& nodes, const In<double>& targets) {
        const Lagrange1D L(values(nodes));
        std::vector<double> out;
        for (py::ssize_t t = 0; t < targets.size(); ++t) {
          const std::vector<double> v = L.values(targets.data()[t]);
          out.insert(out.end(), v.begin(), v.end());
        }
        return array(out, {static_cast<py::ssize_t>(targets.size()), L.size()});
      },
      py::arg("nodes"), py::arg("targets"), "I[t, j] = l_j(targets[t]).");

  // ---- mesh
  py::class_<MacroMesh, std::shared_ptr<MacroMesh>>(m, "MacroMesh")
      .def(py::init(&new_mesh), py::arg("cell_type"), py::arg("coords"), py::arg("nodes"), py::arg("cells"),
           py::arg("material"), py::arg("boundary") = In<int32_t>(std::vector<py::ssize_t>{0, 3}))
      .def_property_readonly("cell_type", [](const MacroMesh& x) { return x.nodes_per_cell() == 4 ? "quad4" : "quad9"; })
      .def_property_readonly("coords", [](const MacroMesh& x) { return x.axisymmetric() ? "axisymmetric" : "cartesian"; })
      .def_property_readonly("n_nodes", &MacroMesh::n_nodes)
      .def_property_readonly("n_cells", &MacroMesh::n_cells)
      .def_property_readonly("n_edges", &MacroMesh::n_edges)
      .def_property_readonly("nodes", [](const MacroMesh& x) { return array(x.nodes(), {x.n_nodes(), 2}); })
      .def_property_readonly("cells", [](const MacroMesh& x) { return array(x.cells(), {x.n_cells(), x.nodes_per_cell()}); })
      .def_property_readonly("material", [](const MacroMesh& x) {
        return array(std::vector<int32_t>(x.materials().begin(), x.materials().end()));
      })
      .def_property_readonly("boundary", [](const MacroMesh& x) {
        std::vector<int32_t> v;
        for (const BoundaryFace& f : x.boundary()) v.insert(v.end(), {f.cell, f.face, f.tag});
        return array(v, {static_cast<py::ssize_t>(x.boundary().size()), 3});
      })
      .def("map", &MacroMesh::map, py::arg("cell"), py::arg("xi"), py::arg("eta"))
      .def("jacobian", &MacroMesh::jacobian, py::arg("cell"), py::arg("xi"), py::arg("eta"))
      .def("__repr__", [](const MacroMesh& x) {
        return "<MacroMesh " + std::to_string(x.n_cells()) + (x.nodes_per_cell() == 4 ? " quad4" : " quad9") +
               " cells, " + (x.axisymmetric() ? "axisymmetric" : "cartesian") + ">";
      });

  py::class_<Forest>(m, "Forest")
      .def_static("uniform", [](std::shared_ptr<MacroMesh> macro, int level) { return Forest::uniform(macro, level); },
                  py::arg("macro"), py::arg("level"))
      .def_property_readonly("n_leaves", &Forest::n_leaves)
      .def_property_readonly(
          "leaves",
          [](const Forest& f) {
            std::vector<int32_t> v;
            v.reserve(4 * static_cast<std::size_t>(f.n_leaves()));
            for (int32_t l = 0; l < f.n_leaves(); ++l) {
              const Leaf& x = f.leaf(l);
              v.insert(v.end(), {x.tree, x.level, x.i, x.j});
            }
            return array(v, {f.n_leaves(), 4});
          },
          "(n_leaves, 4) array of (tree, level, i, j).")
      .def("leaf_box", &Forest::leaf_box, py::arg("leaf"))
      .def_property_readonly("max_level", &Forest::max_level)
      .def(
          "face_neighbor",
          [](const Forest& f, int32_t l, int face) {
            const Neighbor nb = f.neighbor(l, face);
            return py::make_tuple(nb.leaf, nb.face, nb.level_diff);
          },
          py::arg("leaf"), py::arg("face"), "(neighbour leaf or -1, its face, level difference).")
      .def(
          "leaf_corners",
          [](const Forest& f, int32_t l) {
            std::vector<double> v;
            for (auto [xi, eta] : {std::pair{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}}) {
              const auto p = leaf_map(f, l, xi, eta);
              v.insert(v.end(), p.begin(), p.end());
            }
            return array(v, {4, 2});
          },
          py::arg("leaf"), "(4, 2) physical corner coordinates of a leaf.");

  // ---- point maps and geometry
  py::class_<DofMap>(m, "DofMap")
      .def_static(
          "build",
          [](const Forest& f, int ngll) {
            std::vector<int32_t> all(static_cast<std::size_t>(f.n_leaves()));
            for (int32_t l = 0; l < f.n_leaves(); ++l) all[l] = l;
            return DofMap::build(f, ngll, all);
          },
          py::arg("forest"), py::arg("ngll"))
      .def_property_readonly("ngll", [](const DofMap& d) { return d.ngll; })
      .def_property_readonly("n_elements", [](const DofMap& d) { return d.n_elements; })
      .def_property_readonly("n_global", [](const DofMap& d) { return d.n_global; })
      .def_property_readonly(
          "local_to_global", [](const DofMap& d) { return array(d.l2g, {d.n_elements, d.ngll, d.ngll}); },
          "(n_elements, ngll, ngll) global index of each local point, [e, j, i].")
      .def_property_readonly("multiplicity", [](const DofMap& d) { return array(d.multiplicity()); })
      .def("gather", &dof_gather, py::arg("global_field"))
      .def("assemble", &dof_assemble, py::arg("local_field"));

  py::class_<Geometry>(m, "Geometry")
      .def_readonly("ngll", &Geometry::ngll)
      .def_readonly("n_elements", &Geometry::n_elements)
      .def_property_readonly("kind", [](const Geometry& g) {
        static const char* names[] = {"rect", "affine", "curved", "curved"};
        std::vector<std::string> v;
        for (Kind k : g.kind) v.push_back(names[static_cast<int>(k)]);
        return v;
      })
      .def_property_readonly("x", [](const Geometry& g) { return array(g.x, {g.n_elements, g.ngll, g.ngll}); })
      .def_property_readonly("z", [](const Geometry& g) { return array(g.z, {g.n_elements, g.ngll, g.ngll}); })
      .def_property_readonly("jac", [](const Geometry& g) { return array(g.jac, {g.n_elements, g.ngll, g.ngll}); })
      .def_property_readonly("xi_x", [](const Geometry& g) { return array(g.xi_x, {g.n_elements, g.ngll, g.ngll}); })
      .def_property_readonly("xi_z", [](const Geometry& g) { return array(g.xi_z, {g.n_elements, g.ngll, g.ngll}); })
      .def_property_readonly("eta_x", [](const Geometry& g) { return array(g.eta_x, {g.n_elements, g.ngll, g.ngll}); })
      .def_property_readonly("eta_z", [](const Geometry& g) { return array(g.eta_z, {g.n_elements, g.ngll, g.ngll}); });

  m.def(
      "global_coordinates",
      [](const DofMap& d, const Geometry& g) {
        INSEM_REQUIRE(d.n_elements == g.n_elements && d.ngll == g.ngll,
                      "dof map and geometry do not describe the same elements");
        const std::size_t n = static_cast<std::size_t>(d.n_global);
        std::vector<double> xz(2 * n);
        double worst = 0.0;
        for (int32_t p = 0; p < d.n_global; ++p) {
          const std::size_t s0 = static_cast<std::size_t>(d.primary[p]);
          xz[p] = g.x[s0];
          xz[n + p] = g.z[s0];
          for (int32_t q = d.t_offsets[p]; q < d.t_offsets[p + 1]; ++q) {
            const std::size_t s = static_cast<std::size_t>(d.t_slots[q]);
            if (d.l2g[s] != p) continue;  // a constrained slot's contribution
            worst = std::max({worst, std::abs(g.x[s] - xz[p]), std::abs(g.z[s] - xz[n + p])});
          }
        }
        return py::make_tuple(array(xz, {2, d.n_global}), worst);
      },
      py::arg("dofs"), py::arg("geometry"), "(coordinates (2, n_global), max mismatch between element-local copies).");

  // ---- problem description
  py::class_<Material>(m, "Material")
      .def_static(
          "acoustic",
          [](double vp, double rho) {
            Material x{Medium::Acoustic, vp, 0.0, rho};
            x.validate();
            return x;
          },
          py::arg("vp"), py::arg("rho"))
      .def_static(
          "elastic",
          [](double vp, double vs, double rho) {
            Material x{Medium::Elastic, vp, vs, rho};
            x.validate();
            return x;
          },
          py::arg("vp"), py::arg("vs"), py::arg("rho"))
      .def_property_readonly("medium", [](const Material& x) { return medium_name(x.medium); })
      .def_readonly("vp", &Material::vp)
      .def_readonly("vs", &Material::vs)
      .def_readonly("rho", &Material::rho)
      .def("__repr__", &material_repr);

  py::class_<Stf>(m, "SourceTimeFunction")
      .def_property_readonly("kind", [](const Stf& s) { return stf_name(s.type); })
      .def_readonly("f0", &Stf::f0)
      .def_readonly("t0", &Stf::t0)
      .def_readonly("amplitude", &Stf::amplitude)
      .def("__call__", [](const Stf& s, double t) { return s(t); }, py::arg("t"))
      .def(
          "__call__",
          [](const Stf& s, const In<double>& t) {
            py::array_t<double> out(t.size());
            for (py::ssize_t i = 0; i < t.size(); ++i) out.mutable_data()[i] = s(t.data()[i]);
            return out;
          },
          py::arg("t"));
  auto delay = [](py::object t0) { return t0.is_none() ? -1.0 : t0.cast<double>(); };
  for (auto [name, type] : {std::pair{"ricker", int(Stf::Ricker)}, {"gaussian", int(Stf::Gaussian)},
                            {"gaussian_derivative", int(Stf::GaussianDerivative)}}) {
    const int t = type;
    m.def(
        name, [t, delay](double f0, py::object t0, double amp) { return Stf::make(t, f0, delay(t0), amp); },
        py::arg("f0"), py::arg("t0") = py::none(), py::arg("amplitude") = 1.0,
        t == Stf::Ricker ? "Ricker wavelet with dominant frequency f0 (default delay t0 = 1.2 / f0)." : "");
  }
  m.def(
      "sampled_stf",
      [](const In<double>& samples, double dt, double amp) {
        INSEM_REQUIRE(dt > 0.0, "sample interval must be positive");
        INSEM_REQUIRE(samples.size() > 0, "sampled source time function needs samples");
        Stf s;
        s.type = Stf::Samples;
        s.samples = values(samples);
        s.sample_dt = dt;
        s.amplitude = amp;
        return s;
      },
      py::arg("samples"), py::arg("dt"), py::arg("amplitude") = 1.0,
      "Source time function from samples at t = i * dt (linear interpolation).");

  auto source = [](int type, double x, double z, std::array<double, 3> c, const Stf& stf) {
    Source s;
    s.type = type;
    s.position = {x, z};
    s.components = c;
    s.stf = stf;
    return s;
  };
  py::class_<Source>(m, "Source")
      .def_static(
          "force",
          [source](double x, double z, double fx, double fz, const Stf& stf) { return source(Source::Force, x, z, {fx, fz, 0.0}, stf); },
          py::arg("x"), py::arg("z"), py::arg("fx"), py::arg("fz"), py::arg("stf"))
      .def_static(
          "moment_tensor",
          [source](double x, double z, double mxx, double mzz, double mxz, const Stf& stf) {
            return source(Source::MomentTensor, x, z, {mxx, mzz, mxz}, stf);
          },
          py::arg("x"), py::arg("z"), py::arg("mxx"), py::arg("mzz"), py::arg("mxz"), py::arg("stf"))
      .def_static(
          "monopole", [source](double x, double z, const Stf& stf) { return source(Source::Monopole, x, z, {1.0, 0.0, 0.0}, stf); },
          py::arg("x"), py::arg("z"), py::arg("stf"),
          "Acoustic volume-injection source; stf is d^2Q/dt^2 (m^3/s^2 per unit length in 2D).")
      .def_static(
          "dipole",
          [source](double x, double z, double fx, double fz, const Stf& stf) { return source(Source::Dipole, x, z, {fx, fz, 0.0}, stf); },
          py::arg("x"), py::arg("z"), py::arg("fx"), py::arg("fz"), py::arg("stf"),
          "Acoustic point force (pressure dipole) of direction (fx, fz) in N; stf is the force time function.")
      .def_property_readonly("kind", [](const Source& s) { return s.type_name(); })
      .def_readonly("position", &Source::position)
      .def_readonly("components", &Source::components)
      .def_readonly("stf", &Source::stf);

  py::class_<Receiver>(m, "Receiver")
      .def_readonly("name", &Receiver::name)
      .def_readonly("position", &Receiver::position)
      .def_readonly("quantities", &Receiver::quantities)
      .def_readonly("medium", &Receiver::medium);

  py::class_<TimeConfig>(m, "TimeConfig")
      .def_readonly("dt", &TimeConfig::dt)
      .def_readonly("safety", &TimeConfig::safety)
      .def_readonly("n_steps", &TimeConfig::n_steps)
      .def_readonly("duration", &TimeConfig::duration)
      .def_readonly("record_every", &TimeConfig::record_every)
      .def_readonly("energy_every", &TimeConfig::energy_every)
      .def_readonly("snapshot_every", &TimeConfig::snapshot_every)
      .def_readonly("snapshot_quantities", &TimeConfig::snapshot_quantities);

  // ---- solver (the time loop on the GPU)
  py::class_<Solver>(m, "Solver")
      .def_property_readonly("dt", &Solver::dt)
      .def_property_readonly("stable_dt", &Solver::stable_dt)
      .def_property_readonly("n_steps", &Solver::n_steps)
      .def_property_readonly("step", &Solver::step)
      .def_property_readonly("time", &Solver::time)
      .def("advance", &Solver::advance, py::arg("n_steps"), py::call_guard<py::gil_scoped_release>())
      .def("run", &Solver::run, py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("media", [](const Solver& s) {
        std::vector<std::string> v;
        for (const auto& d : s.image().domains) v.push_back(medium_name(d.medium));
        return v;
      })
      .def(
          "coordinates",
          [](const Solver& s, const std::string& medium) {
            const auto& d = solver_domain(s, medium);
            return array(d.coords, {2, d.dofs.n_global});
          },
          py::arg("medium"), "(2, n_points) coordinates of the global points of a medium.")
      .def(
          "state",
          [](const Solver& s, const std::string& medium) {
            const auto& d = solver_domain(s, medium);
            const int di = s.domain_index(medium_arg(medium));
            auto f = [&](int which) { return array(s.field(di, which), {d.ncomp, d.dofs.n_global}); };
            return py::make_tuple(f(INSEM_FIELD_U), f(INSEM_FIELD_V), f(INSEM_FIELD_A));
          },
          py::arg("medium"),
          "(field, first, second time derivative), each (ncomp, n_points): displacement (elastic) or pressure "
          "(acoustic). Azimuthal order 1 elastic components are (U+, U-, U_z), U+- = (U_r +- U_theta) / sqrt 2.")
      .def(
          "set_state",
          [](Solver& s, const std::string& medium, const In<double>& u, const In<double>& v) {
            s.set_state(medium_arg(medium), values(u), values(v));
          },
          py::arg("medium"), py::arg("u"), py::arg("v"))
      .def(
          "trace",
          [](const Solver& s, const std::string& receiver, const std::string& quantity) {
            for (const Solver::Trace& t : s.traces())
              if (t.receiver == receiver && t.quantity == quantity)
                return array(t.data, {static_cast<py::ssize_t>(t.data.size() / t.ncomp), t.ncomp});
            throw Error("no trace '" + quantity + "' for receiver '" + receiver + "'");
          },
          py::arg("receiver"), py::arg("quantity"), "(n_samples, ncomp) recorded trace.")
      .def_property_readonly("times", [](const Solver& s) { return array(s.sample_times()); })
      .def_property_readonly(
          "energy",
          [](const Solver& s) {
            std::vector<double> e;
            const auto rows = s.energy();
            for (const auto& r : rows) e.insert(e.end(), r.begin(), r.end());
            return array(e, {static_cast<py::ssize_t>(rows.size()), 3});
          },
          "(n, 3) energy monitor samples: time, kinetic, potential.")
      .def(
          "snapshot",
          [](const Solver& s, const std::string& medium, const std::string& quantity) {
            const Medium md = medium_arg(medium);
            for (const Solver::Snapshot& sn : s.snapshots()) {
              if (sn.medium != md || sn.quantity != quantity) continue;
              const py::ssize_t ns = static_cast<py::ssize_t>(s.snapshot_times().size());
              const py::ssize_t n = ns ? static_cast<py::ssize_t>(sn.data.size()) / (ns * sn.ncomp) : 0;
              return array(sn.data, {ns, sn.ncomp, n});
            }
            throw Error("no " + quantity + " snapshots for the " + medium + " medium");
          },
          py::arg("medium"), py::arg("quantity"), "(n_snapshots, ncomp, n_points) snapshot data.")
      .def(
          "snapshot_points",
          [](const Solver& s, const std::string& medium) {
            const std::vector<double> p = s.snapshot_points(medium_arg(medium));
            return array(p, {2, static_cast<py::ssize_t>(p.size() / 2)});
          },
          py::arg("medium"), "(2, n_points) coordinates of the snapshot points.")
      .def(
          "snapshot_cells",
          [](const Solver& s, const std::string& medium) {
            const std::vector<int32_t> c = s.snapshot_cells(medium_arg(medium));
            return array(c, {static_cast<py::ssize_t>(c.size() / 4), 4});
          },
          py::arg("medium"), "(n_cells, 4) sub-quad connectivity of the snapshot points.")
      .def_property_readonly("snapshot_times", [](const Solver& s) { return array(s.snapshot_times()); })
      .def("write_results", [](const Solver& s, const std::string& path) { write_results(s, path); }, py::arg("path"));

  // ---- reading case and results files without h5py (not in the reference)
  m.def(
      "_read_dataset",
      [](const std::string& path, const std::string& name) {
        const h5::File f(path, h5::Open::Read);
        std::vector<std::size_t> shape;
        const std::vector<double> v = f.get<double>(name, &shape);
        return array(v, std::vector<py::ssize_t>(shape.begin(), shape.end()));
      },
      py::arg("path"), py::arg("name"), "A dataset of an HDF5 file as a float64 array of its shape.");
  m.def(
      "_read_attribute",
      [](const std::string& path, const std::string& object, const std::string& name) -> py::object {
        const h5::File f(path, h5::Open::Read);
        switch (f.attr_class(object, name)) {
          case 's': return py::str(f.get_string(object, name));
          case 'i': return py::int_(f.get_int(object, name));
          case 'f': return py::float_(f.get_double(object, name));
          default: throw Error("attribute '" + name + "' of '" + object + "' is neither text nor a number");
        }
      },
      py::arg("path"), py::arg("object"), py::arg("name"), "A scalar attribute of an HDF5 file.");
  m.def(
      "_has_object", [](const std::string& path, const std::string& name) { return h5::File(path, h5::Open::Read).has(name); },
      py::arg("path"), py::arg("name"));

  // ---- simulation
  py::class_<Discretization>(m, "Discretization")
      .def(py::init([](int order, int uniform_level, double f_max, double ppw, int max_level,
                       const std::vector<RegionTuple>& regions, int azimuthal_order) {
             Discretization d;
             d.order = order;
             d.uniform_level = uniform_level;
             d.f_max = f_max;
             d.points_per_wavelength = ppw;
             d.max_level = max_level;
             d.azimuthal_order = azimuthal_order;
             for (const auto& [x0, x1, z0, z1, level] : regions) d.regions.push_back({x0, x1, z0, z1, level});
             return d;
           }),
           py::arg("order") = 4, py::arg("uniform_level") = 0, py::arg("f_max") = 0.0,
           py::arg("points_per_wavelength") = 5.0, py::arg("max_level") = 12,
           py::arg("regions") = std::vector<RegionTuple>{}, py::arg("azimuthal_order") = 0,
           "Element order and refinement: uniform levels, then adaptive refinement to points_per_wavelength "
           "at f_max (0 = off) and to the levels of regions [(xmin, xmax, zmin, zmax, level), ...], capped at "
           "max_level; then 2:1 balance. azimuthal_order (axisymmetric meshes): 0 monopole, 1 dipole fields.")
      .def_readonly("order", &Discretization::order)
      .def_readonly("uniform_level", &Discretization::uniform_level)
      .def_readonly("f_max", &Discretization::f_max)
      .def_readonly("points_per_wavelength", &Discretization::points_per_wavelength)
      .def_readonly("max_level", &Discretization::max_level)
      .def_readonly("azimuthal_order", &Discretization::azimuthal_order)
      .def_property_readonly("regions", [](const Discretization& d) {
        std::vector<RegionTuple> v;
        for (const Region& r : d.regions) v.emplace_back(r.xmin, r.xmax, r.zmin, r.zmax, r.level);
        return v;
      });

  py::class_<Sim>(m, "Simulation")
      .def(py::init([](std::shared_ptr<MacroMesh> mesh, const Discretization& disc) {
             return Sim{std::make_unique<Model>(mesh, disc), nullptr};
           }),
           py::arg("mesh"), py::arg("discretization") = Discretization{})
      .def("setup", [](Sim& s) { s.model->setup(); }, py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("is_setup", [](const Sim& s) { return s.model->is_setup(); })
      .def_property_readonly("mesh", [](const Sim& s) { return std::const_pointer_cast<MacroMesh>(s.model->mesh_ptr()); })
      .def_property_readonly("discretization", [](const Sim& s) { return s.model->discretization(); })
      .def_property_readonly("forest", [](const Sim& s) -> const Forest& { return s.model->forest(); },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("dofs", [](const Sim& s) -> const DofMap& { return s.model->dofs(); },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("geometry", [](const Sim& s) -> const Geometry& { return s.model->geometry(); },
                             py::return_value_policy::reference_internal)
      .def_property_readonly("gll", [](const Sim& s) { return rule_tuple(s.model->gll()); })
      .def("geometry_kind_counts", [](const Sim& s) { return s.model->kind_counts(); })
      .def(
          "set_material",
          [](Sim& s, int region, const Material& mat) {
            s.model->set_material(region, mat);
            s.solver.reset();
          },
          py::arg("region"), py::arg("material"))
      .def(
          "set_boundary",
          [](Sim& s, int tag, const std::string& type) {
            s.model->set_boundary(tag, boundary_from_name(type));
            s.solver.reset();
          },
          py::arg("tag"), py::arg("type"), "type: 'free', 'rigid' or 'absorbing'.")
      .def(
          "add_source",
          [](Sim& s, const Source& src) {
            s.model->add_source(src);
            s.solver.reset();
          },
          py::arg("source"))
      .def("clear_sources", [](Sim& s) {
        s.model->clear_sources();
        s.solver.reset();
      })
      .def(
          "add_receiver",
          [](Sim& s, const std::string& name, double x, double z, std::vector<std::string> quantities,
             const std::string& medium) {
            s.model->add_receiver({name, {x, z}, std::move(quantities), medium});
            s.solver.reset();
          },
          py::arg("name"), py::arg("x"), py::arg("z"), py::arg("quantities"), py::arg("medium") = "",
          "medium ('acoustic' / 'elastic') picks the side of a receiver on a fluid-solid interface.")
      .def("clear_receivers", [](Sim& s) {
        s.model->clear_receivers();
        s.solver.reset();
      })
      .def(
          "set_time",
          [](Sim& s, int n_steps, double duration, double dt, double safety, int record_every, int energy_every,
             int snapshot_every, std::vector<std::string> snapshot_quantities) {
            TimeConfig t;
            t.n_steps = n_steps;
            t.duration = duration;
            t.dt = dt;
            t.safety = safety;
            t.record_every = record_every;
            t.energy_every = energy_every;
            t.snapshot_every = snapshot_every;
            t.snapshot_quantities = std::move(snapshot_quantities);
            s.model->set_time(t);
            s.solver.reset();
          },
          py::arg("n_steps") = 0, py::arg("duration") = 0.0, py::arg("dt") = 0.0, py::arg("safety") = 0.8,
          py::arg("record_every") = 1, py::arg("energy_every") = 0, py::arg("snapshot_every") = 0,
          py::arg("snapshot_quantities") = std::vector<std::string>{"velocity", "pressure"},
          "Time settings. dt = 0 picks safety * the estimated stability limit; n_steps = 0 derives it from duration.")
      .def_property_readonly("materials", [](const Sim& s) { return s.model->materials(); })
      .def_property_readonly("boundaries", [](const Sim& s) {
        std::map<int, std::string> v;
        for (const auto& [tag, b] : s.model->boundaries()) v[tag] = boundary_name(b);
        return v;
      })
      .def_property_readonly("sources", [](const Sim& s) { return s.model->sources(); })
      .def_property_readonly("receivers", [](const Sim& s) { return s.model->receivers(); })
      .def_property_readonly("time", [](const Sim& s) { return s.model->time(); })
      .def_property_readonly("solver", [](Sim& s) -> Solver& { return s.get_solver(); },
                             py::return_value_policy::reference_internal, "Time-stepping state, built on first access.")
      .def("run", [](Sim& s) { s.get_solver().run(); }, py::call_guard<py::gil_scoped_release>())
      .def("write_results", [](Sim& s, const std::string& path) { write_results(s.get_solver(), path); }, py::arg("path"))
      .def("save", [](const Sim& s, const std::string& path) { save_case(*s.model, path); }, py::arg("path"))
      .def_static("load", [](const std::string& path) { return Sim{load_case(path), nullptr}; }, py::arg("path"));
}
