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

  insem_hip_ctx* context() const { return ctx_; }

 private:
  std::unique_ptr<SetupImage> img_;
  insem_hip_ctx* ctx_ = nullptr;
  double dt_ = 0.0, stable_dt_ = 0.0;
  int n_steps_ = 0;
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
