// Many independent models as one setup: the disjoint union of their
// domains, blocks, points, boundary data, coupling, sources and receivers.
//
// The union is what the batched backend steps: one element-force launch per
// element group and one tail per domain cover every model at once, instead
// of one set of launches per model.  Models keep their own time step and
// step count; every per-point computation only touches its own model's
// data, and each point's assembly sum keeps its model's slot order, so a
// model's results in the union are bitwise its results alone.
#pragma once

#include <memory>
#include <vector>

#include "model.hpp"
#include "solver.hpp"

namespace insem_amd {

struct BatchImage {
  std::unique_ptr<SetupImage> image;            // the union
  std::vector<std::vector<int32_t>> point_model;  // per domain: model of each global point
  std::vector<int32_t> receiver_model;           // model of each union receiver
  std::vector<int32_t> source_model;             // model of each union source
  std::vector<int32_t> receiver_offset;          // first union receiver of each model
  int n_models = 0;
};

// All models must share the polynomial order, the coordinate system and
// azimuthal order, the media present and record_every.
BatchImage union_of(const std::vector<const SetupImage*>& models);

// One model's power iteration (lambda_max of M^-1 K, as a Solver of the
// model alone runs it) queued on a short-lived context of that model: the
// constructor uploads the model and queues the iterations without waiting,
// ready() asks whether they have finished, lambda() waits for the result
// and frees the context.  The model is copied to the device by the
// constructor and not read afterwards.
// The host half of a power iteration's start (insem_hip_stage: the
// context laid out for a reserved block, and the start vectors), made on any
// thread without device calls; PowerJob(StagedPower&) finishes it.
struct StagedPower {
  StagedPower(const SetupImage& model, void* block);
  ~StagedPower();  // a context never realized gives its block back
  StagedPower(const StagedPower&) = delete;
  StagedPower& operator=(const StagedPower&) = delete;
  insem_hip_ctx* ctx = nullptr;
  std::vector<float> start;
};

class PowerJob {
 public:
  explicit PowerJob(const SetupImage& model);
  // Realizes a staged context (device calls) and starts its iteration; the
  // staged object keeps nothing afterwards.
  explicit PowerJob(StagedPower& staged);
  ~PowerJob();
  PowerJob(const PowerJob&) = delete;
  PowerJob& operator=(const PowerJob&) = delete;
  bool ready() const;
  double lambda();
  // Waits for lambda_max and hands the context over (for a Solver of the
  // model: the context is the one that Solver would build); the job keeps
  // no context afterwards.
  insem_hip_ctx* take_context(double* lambda);

 private:
  insem_hip_ctx* ctx_ = nullptr;
  double lambda_ = 0.0;
};

// Many models stepped as one union on the device (insem_hip_create_batch).
// Each model's Δt comes from its own power iteration (a short-lived
// single-model context), as a Solver of that model alone would get it; the
// second constructor takes those lambda_max values already computed (by
// PowerJobs started earlier), so that it makes no power iteration itself.
class BatchSolver {
 public:
  explicit BatchSolver(std::vector<std::unique_ptr<SetupImage>> models);
  BatchSolver(std::vector<std::unique_ptr<SetupImage>> models, const std::vector<double>& lambdas);
  ~BatchSolver();
  BatchSolver(const BatchSolver&) = delete;
  BatchSolver& operator=(const BatchSolver&) = delete;

  int n_models() const { return static_cast<int>(models_.size()); }
  // m: the model's place in the list the batch was made from.
  const SetupImage& model(int m) const { return *models_.at(slot(m)); }
  double dt(int m) const { return time_.at(slot(m)).dt; }
  double stable_dt(int m) const { return time_.at(slot(m)).stable_dt; }
  int n_steps(int m) const { return time_.at(slot(m)).n_steps; }
  int max_steps() const { return max_steps_; }
  int step() const;
  void advance(int n);
  void run() { advance(max_steps_ - step()); }
  // Samples of model m so far (its own record cadence, up to its n_steps).
  int n_samples(int m) const;
  // Model m's traces, as Solver::traces of that model alone.
  std::vector<Solver::Trace> traces(int m) const;
  // Steps the device has completed (does not wait), and model m's traces
  // read without waiting for the steps still queued (its last step must be
  // done on the device).
  int device_step() const;
  std::vector<Solver::Trace> traces_done(int m) const;
  insem_hip_ctx* context() const { return ctx_; }
  int64_t footprint() const;

 private:
  std::vector<Solver::Trace> model_traces(const std::vector<double>& all, int ns, int m) const;
  std::size_t trace_values(int ns) const;
  void start(const std::vector<double>& lambdas);
  int slot(int m) const { return slot_.at(m); }

  // The models in union order: by step count, longest first, so that the
  // work of the models still running is a leading part of every launch
  // (the backend then launches only that part); slot_[m] is the union
  // place of the caller's model m.
  std::vector<std::unique_ptr<SetupImage>> models_;
  std::vector<int> slot_;
  BatchImage union_;
  std::vector<ModelTime> time_;
  int max_steps_ = 0;
  insem_hip_ctx* ctx_ = nullptr;
};

}  // namespace insem_amd
