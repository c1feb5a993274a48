#include "solver.hpp"

#include <algorithm>
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
    set_up_time(device_lambda_max(ctx_, *img_));
  } catch (...) {
    insem_hip_destroy(ctx_);
    ctx_ = nullptr;
    throw;
  }
}

Solver::Solver(std::unique_ptr<SetupImage> image, insem_hip_ctx* ctx, double lambda)
    : img_(std::move(image)), ctx_(ctx) {
  INSEM_REQUIRE(img_ && ctx_, "solver needs a setup and its context");
  set_up_time(lambda);
}

void Solver::set_up_time(double lambda) {
  const ModelTime m = model_time(img_->time, lambda);
  stable_dt_ = m.stable_dt;
  dt_ = m.dt;
  n_steps_ = m.n_steps;
  const std::vector<float> amps = img_->source_amplitudes(dt_, n_steps_);
  check_hip(insem_hip_set_time(ctx_, dt_, n_steps_, amps.data()));
  if (img_->time.snapshot_every > 0)
    for (const std::string& q : img_->time.snapshot_quantities)
      for (const auto& d : img_->domains) {
        const bool has = d.medium == Medium::Elastic
                             ? (q == "displacement" || q == "velocity" || q == "acceleration")
                             : q == "pressure";
        if (has) snapshots_.push_back({d.medium, q, d.ncomp, {}});
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

// With snapshots the device runs to each snapshot step in turn (an advance
// call that ends on a step leaves the fields as the reference's step does;
// splitting a run into calls changes no value).
void Solver::advance(int n) {
  const int every = img_->time.snapshot_every;
  if (every <= 0) {
    check_hip(insem_hip_advance(ctx_, n));
    return;
  }
  if (!started_) {
    check_hip(insem_hip_advance(ctx_, 0));  // initialisation only
    record_snapshot();
    started_ = true;
  }
  int s = step();
  const int target = std::min(n_steps_, s + std::max(n, 0));
  while (s < target) {
    const int next = std::min(target, (s / every + 1) * every);
    check_hip(insem_hip_advance(ctx_, next - s));
    s = next;
    if (s % every == 0) record_snapshot();
  }
}

const SetupImage::Domain& Solver::domain_of(Medium m) const {
  const int di = img_->domain_index(m);
  INSEM_REQUIRE(di >= 0, "the model has no " << medium_name(m) << " region");
  return img_->domains[di];
}

// One snapshot: the device field gathered to the element points as
// DofMap::gather does it (dof_map.hpp:94-120: constrained slots sum
// float(weight) * value in row order, in float).
void Solver::record_snapshot() {
  snapshot_times_.push_back(time());
  for (Snapshot& sn : snapshots_) {
    const int di = img_->domain_index(sn.medium);
    const auto& d = img_->domains[di];
    const auto& M = d.dofs;
    const int which = sn.quantity == "velocity" ? INSEM_FIELD_V : sn.quantity == "acceleration" ? INSEM_FIELD_A
                                                                                                   : INSEM_FIELD_U;
    const std::size_t ng = static_cast<std::size_t>(M.n_global), npe = static_cast<std::size_t>(M.npe);
    const std::size_t n = static_cast<std::size_t>(M.n_elements) * npe;
    std::vector<float> g(static_cast<std::size_t>(d.ncomp) * ng);
    check_hip(insem_hip_get_field(ctx_, di, which, g.data()));
    const std::size_t base = sn.data.size();
    sn.data.resize(base + static_cast<std::size_t>(d.ncomp) * n);
    float* out = &sn.data[base];
    for (int c = 0; c < d.ncomp; ++c) {
      const float* gc = &g[c * ng];
      float* oc = out + c * n;
      for (std::size_t s = 0; s < n; ++s) {
        const int32_t m = M.l2g[s];
        if (m >= 0) {
          oc[s] = gc[m];
        } else {
          const int32_t r = -1 - m;
          float acc = 0.0f;
          for (int32_t q = M.row_offsets[r]; q < M.row_offsets[r + 1]; ++q)
            acc += static_cast<float>(M.row_weights[q]) * gc[M.row_dofs[q]];
          oc[s] = acc;
        }
      }
    }
    if (sn.medium == Medium::Elastic && d.azimuthal_order == 1) {
      // (U+, U-) -> (U_r, U_theta) = ((U+ + U-), (U+ - U-)) / sqrt 2, in double
      const double h = 0.70710678118654752440;
      for (std::size_t s = 0; s < n; ++s) {
        const double up = out[s], um = out[n + s];
        out[s] = static_cast<float>(h * (up + um));
        out[n + s] = static_cast<float>(h * (up - um));
      }
    }
  }
}

std::vector<double> Solver::snapshot_points(Medium m) const {
  const auto& d = domain_of(m);
  const std::size_t n = d.point_x.size();
  INSEM_REQUIRE(n > 0 || d.dofs.n_elements == 0, "the model takes no snapshots (snapshot_every is 0)");
  std::vector<double> out(2 * n);
  std::copy(d.point_x.begin(), d.point_x.end(), out.begin());
  std::copy(d.point_z.begin(), d.point_z.end(), out.begin() + static_cast<std::ptrdiff_t>(n));
  return out;
}

// Every element's (ngll - 1)^2 sub-quads over its points, counter-clockwise.
std::vector<int32_t> Solver::snapshot_cells(Medium m) const {
  const auto& d = domain_of(m);
  const int ng = d.dofs.ngll, npe = d.dofs.npe;
  std::vector<int32_t> out;
  out.reserve(static_cast<std::size_t>(d.dofs.n_elements) * (ng - 1) * (ng - 1) * 4);
  for (int32_t e = 0; e < d.dofs.n_elements; ++e)
    for (int j = 0; j + 1 < ng; ++j)
      for (int i = 0; i + 1 < ng; ++i) {
        const int32_t p = e * npe + j * ng + i;
        out.insert(out.end(), {p, p + 1, p + ng + 1, p + ng});
      }
  return out;
}

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
