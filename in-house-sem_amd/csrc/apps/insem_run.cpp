// insem-run CASE.h5 [RESULTS.h5]: the reference's command-line runner
// (cpp/apps/insem_run.cpp:23-65) on this backend -- load a case file, set the
// model up, report its size, run every step on the GPU and write the results
// file (default: the case path with "_results.h5" in place of ".h5").
#include <chrono>
#include <cstdio>
#include <exception>
#include <string>

#include "../host/solver.hpp"
#include "../io/files.hpp"

using namespace insem_amd;

namespace {

std::string results_path_for(const std::string& case_path) {
  const std::size_t k = case_path.rfind(".h5");
  return (k == std::string::npos ? case_path : case_path.substr(0, k)) + "_results.h5";
}

int usage(int code) {
  std::fprintf(stderr, "usage: insem-run CASE.h5 [RESULTS.h5]\n");
  return code;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return usage(2);
  const std::string first = argv[1];
  if (first == "-h" || first == "--help") return usage(0);
  if (argc > 3) return usage(2);
  const std::string results = argc == 3 ? argv[2] : results_path_for(first);
  try {
    std::unique_ptr<Model> model = load_case(first);
    model->setup();
    const MacroMesh& mesh = model->mesh();
    std::printf("insem %s (MI355X backend)\n", insem_version());
    std::printf("  case:            %s\n", first.c_str());
    std::printf("  coordinates:     %s\n", mesh.axisymmetric() ? "axisymmetric" : "cartesian");
    std::printf("  macro-cells:     %d (%s)\n", mesh.n_cells(), mesh.nodes_per_cell() == 4 ? "quad4" : "quad9");
    std::printf("  order:           %d\n", model->discretization().order);
    std::printf("  elements:        %d\n", model->forest().n_leaves());
    std::printf("  geometry kinds: ");
    for (const auto& [kind, count] : model->kind_counts()) std::printf(" %s=%d", kind.c_str(), count);
    std::printf("\n");
    Solver solver(model->build());
    for (const auto& d : solver.image().domains)
      std::printf("  %s domain:  %d elements, %d points\n", medium_name(d.medium), d.dofs.n_elements, d.dofs.n_global);
    std::printf("  dt:              %g s (stability limit %g s)\n", solver.dt(), solver.stable_dt());
    std::printf("  steps:           %d\n", solver.n_steps());
    std::fflush(stdout);
    const auto t0 = std::chrono::steady_clock::now();
    solver.run();
    check_hip(insem_hip_synchronize(solver.context()));
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    write_results(solver, results);
    std::printf("  wall time:       %g s\n", seconds);
    std::printf("  results:         %s\n", results.c_str());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "insem-run: %s\n", e.what());
    return 1;
  }
  return 0;
}
