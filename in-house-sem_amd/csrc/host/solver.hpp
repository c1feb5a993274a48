// Host mirror of insem's solver::Solver (reference: cpp/include/insem/solver/solver.hpp:91-153)
// running the time loop through the HIP C-ABI (include/insem_hip.h).
//
// Construction follows Solver::Solver (cpp/src/solver/solver.cpp:91-145):
// the stable time step comes from the power iteration (on the device), dt
// from the safety factor or the user's value (checked against the limit),
// n_steps from the duration.  The device context is owned here and freed in
// the destructor.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/insem_hip.h"
#include "model.hpp"

namespace insem_amd {

class Solver {
 public:
  explicit Solver(std::unique_ptr<SetupImage> image);
  // The solver of a context already created from this image whose power
  // iteration gave lambda (a finished PowerJob's): the same state as the
  // constructor above reaches, without building the context again.  Takes
  // the context over.
  Solver(std::unique_ptr<SetupImage> image, insem_hip_ctx* ctx, double lambda);
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  double dt() const { return dt_; }
  double stable_dt() const { return stable_dt_; }
  int n_steps() const { return n_steps_; }
  int step() const;
  double time() const { return step() * dt_; }

  void advance(int n);
  void run() { advance(n_steps_ - step()); }

  const SetupImage& image() const { return *img_; }
  int domain_index(Medium m) const { return img_->domain_index(m); }
  void set_state(Medium m, const std::vector<double>& u, const std::vector<double>& v);
  // Field of a domain in double, [ncomp][n_global]; which = INSEM_FIELD_*.
  std::vector<double> field(int domain, int which) const;

  struct Trace {
    std::string receiver, quantity;
    int ncomp = 0;
    std::vector<double> data;  // [n_samples][ncomp]
  };
  std::vector<Trace> traces() const;
  std::vector<double> sample_times() const;
  std::vector<std::array<double, 3>> energy() const;

  // Snapshots (solver.cpp:813-876): after initialisation and after every
  // snapshot_every-th step, each listed quantity of each medium that has it,
  // on the element points ([n_snapshots][ncomp][n_elements * npe], elements
  // in the domain's leaf order; m = 1 solids as U_r, U_theta, U_z).
  struct Snapshot {
    Medium medium = Medium::Elastic;
    std::string quantity;
    int ncomp = 0;
    std::vector<float> data;
  };
  const std::vector<Snapshot>& snapshots() const { return snapshots_; }
  const std::vector<double>& snapshot_times() const { return snapshot_times_; }
  std::vector<double> snapshot_points(Medium m) const;   // [2][n_points]
  std::vector<int32_t> snapshot_cells(Medium m) const;   // [n_cells][4] sub-quads

  insem_hip_ctx* context() const { return ctx_; }

 private:
  std::unique_ptr<SetupImage> img_;
  void set_up_time(double lambda);  // dt, n_steps, source amplitudes, snapshots
  void record_snapshot();
  const SetupImage::Domain& domain_of(Medium m) const;

  insem_hip_ctx* ctx_ = nullptr;
  double dt_ = 0.0, stable_dt_ = 0.0;
  int n_steps_ = 0;
  std::vector<Snapshot> snapshots_;
  std::vector<double> snapshot_times_;
  bool started_ = false;  // the step-0 snapshot is taken
};

// Throws insem_amd::Error with the backend's message when rc != 0.
void check_hip(int rc);

// Time step and step count from lambda_max(M^-1 K) and the time settings
// (Solver::Solver, solver.cpp:131-144).
struct ModelTime {
  double stable_dt = 0.0, dt = 0.0;
  int n_steps = 0;
};
ModelTime model_time(const TimeConfig& t, double lambda);
// lambda_max by the device power iteration on a context of the model.
double device_lambda_max(insem_hip_ctx* ctx, const SetupImage& img);

}  // namespace insem_amd
