#include "solver.hpp"

#include <cmath>

#include "error.hpp"

namespace insem_amd {

void check_hip(int rc) {
  if (rc != INSEM_HIP_OK) throw Error(insem_hip_last_error());
}

ModelTime model_time(const TimeConfig& t, double lambda) {
  INSEM_REQUIRE(lambda > 0.0, "could not estimate the stable time step (empty or rigid model?)");
  ModelTime m;
  m.stable_dt = 2.0 / std::sqrt(lambda);
  if (t.dt > 0.0) {
    INSEM_REQUIRE(t.dt <= m.stable_dt, "time step " << t.dt << " s exceeds the estimated stability limit "
                                                   << m.stable_dt << " s");
    m.dt = t.dt;
  } else {
    INSEM_REQUIRE(t.safety > 0.0 && t.safety <= 1.0, "time step safety factor must be in (0, 1]");
    m.dt = t.safety * m.stable_dt;
  }
  if (t.n_steps > 0)
    m.n_steps = t.n_steps;
  else if (t.duration > 0.0)
    m.n_steps = static_cast<int>(std::ceil(t.duration / m.dt - 1e-9));
  return m;
}

double device_lambda_max(insem_hip_ctx* ctx, const SetupImage& img) {
  const std::vector<float> start = img.power_start_vectors();
  double lambda = 0.0;
  check_hip(insem_hip_lambda_max(ctx, start.data(), &lambda));
  return lambda;
}

Solver::Solver(std::unique_ptr<SetupImage> image) : img_(std::move(image)) {
  INSEM_REQUIRE(img_, "solver needs a setup");
  check_hip(insem_hip_create(&img_->view, &ctx_));
  try {
    const ModelTime m = model_time(img_->time, device_lambda_max(ctx_, *img_));
    stable_dt_ = m.stable_dt;
    dt_ = m.dt;
    n_steps_ = m.n_steps;
    const std::vector<float> amps = img_->source_amplitudes(dt_, n_steps_);
    check_hip(insem_hip_set_time(ctx_, dt_, n_steps_, amps.data()));
  } catch (...) {
    insem_hip_destroy(ctx_);
    ctx_ = nullptr;
    throw;
  }
}

Solver::~Solver() {
  if (ctx_) insem_hip_destroy(ctx_);
}

int Solver::step() const {
  int32_t s = 0;
  check_hip(insem_hip_step(ctx_, &s));
  return s;
}

void Solver::advance(int n) { check_hip(insem_hip_advance(ctx_, n)); }

void Solver::set_state(Medium m, const std::vector<double>& u, const std::vector<double>& v) {
  const int di = img_->domain_index(m);
  INSEM_REQUIRE(di >= 0, "the model has no " << medium_name(m) << " region");
  const auto& d = img_->domains[di];
  const std::size_t n = static_cast<std::size_t>(d.ncomp) * d.dofs.n_global;
  INSEM_REQUIRE(u.size() == n && v.size() == n,
                "state arrays must have " << d.ncomp << " x " << d.dofs.n_global << " entries");
  std::vector<float> fu(u.begin(), u.end()), fv(v.begin(), v.end());
  check_hip(insem_hip_set_state(ctx_, di, fu.data(), fv.data()));
}

std::vector<double> Solver::field(int domain, int which) const {
  INSEM_REQUIRE(domain >= 0 && domain < static_cast<int>(img_->domains.size()), "invalid domain " << domain);
  const auto& d = img_->domains[domain];
  std::vector<float> f(static_cast<std::size_t>(d.ncomp) * d.dofs.n_global);
  check_hip(insem_hip_get_field(ctx_, domain, which, f.data()));
  return std::vector<double>(f.begin(), f.end());
}

std::vector<Solver::Trace> Solver::traces() const {
  int32_t ns = 0;
  check_hip(insem_hip_n_samples(ctx_, &ns));
  std::size_t total = 0;
  for (const auto& r : img_->receivers)
    for (int nc : r.ncomp) total += static_cast<std::size_t>(nc) * ns;
  std::vector<double> all(total + 1);
  check_hip(insem_hip_read_traces(ctx_, all.data()));
  std::vector<Trace> out;
  std::size_t o = 0;
  for (const auto& r : img_->receivers)
    for (std::size_t q = 0; q < r.quantities.size(); ++q) {
      Trace t;
      t.receiver = r.name;
      t.quantity = r.quantities[q];
      t.ncomp = r.ncomp[q];
      const std::size_t n = static_cast<std::size_t>(t.ncomp) * ns;
      t.data.assign(all.begin() + o, all.begin() + o + n);
      o += n;
      out.push_back(std::move(t));
    }
  return out;
}

std::vector<double> Solver::sample_times() const {
  int32_t ns = 0;
  check_hip(insem_hip_n_samples(ctx_, &ns));
  std::vector<double> t(ns);
  for (int i = 0; i < ns; ++i) t[i] = static_cast<double>(i * img_->time.record_every) * dt_;
  return t;
}

std::vector<std::array<double, 3>> Solver::energy() const {
  int32_t n = 0;
  check_hip(insem_hip_n_energy(ctx_, &n));
  std::vector<double> e(3 * static_cast<std::size_t>(n) + 1);
  check_hip(insem_hip_read_energy(ctx_, e.data()));
  std::vector<std::array<double, 3>> out(n);
  for (int i = 0; i < n; ++i) out[i] = {e[3 * i], e[3 * i + 1], e[3 * i + 2]};
  return out;
}

}  // namespace insem_amd
