// Case and results files (HDF5) of the reference's formats, so that the
// reference's case.h5 files load here and results files read alike.
//
// Case file "insem-case" version 1 (cpp/src/sim/simulation.cpp:233-449):
// /mesh (nodes, cells, material, boundary; cell_type and coords
// attributes), /discretization, /materials/<region>, /boundary_conditions,
// /sources/<i>, /receivers/<i>, /time.  Results file "insem-results"
// version 1 (cpp/src/solver/solver.cpp:965-1016): /time, traces under
// /receivers/<name>/<quantity>, /snapshots and /energy.
#pragma once

#include <memory>
#include <string>

#include "../host/model.hpp"
#include "../host/solver.hpp"

namespace insem_amd {

void save_case(const Model& model, const std::string& path);
// The model a case file describes (not set up).  Description groups other
// than /mesh and /discretization are optional.
std::unique_ptr<Model> load_case(const std::string& path);

// Everything the solver recorded so far.
void write_results(const Solver& solver, const std::string& path);

}  // namespace insem_amd
