#include "batch.hpp"

#include <algorithm>
#include <string>

#include "error.hpp"

namespace insem_amd {
namespace {

template <class T>
void append(std::vector<T>& dst, const std::vector<T>& src) {
  dst.insert(dst.end(), src.begin(), src.end());
}

template <class T>
void append_shifted(std::vector<T>& dst, const std::vector<T>& src, T shift) {
  for (T v : src) dst.push_back(v + shift);
}

}  // namespace

BatchImage union_of(const std::vector<const SetupImage*>& models) {
  INSEM_REQUIRE(!models.empty(), "a batch needs at least one model");
  const SetupImage& first = *models.front();
  for (const SetupImage* m : models) {
    INSEM_REQUIRE(m->ngll == first.ngll, "batched models must share the polynomial order");
    INSEM_REQUIRE(m->axisym == first.axisym && m->azimuthal_order == first.azimuthal_order,
                  "batched models must share the coordinate system and azimuthal order");
    INSEM_REQUIRE(m->domains.size() == first.domains.size(), "batched models must have the same media");
    for (std::size_t di = 0; di < m->domains.size(); ++di)
      INSEM_REQUIRE(m->domains[di].medium == first.domains[di].medium, "batched models must have the same media");
    INSEM_REQUIRE(m->time.record_every == first.time.record_every, "batched models must share record_every");
  }
  const int npe = first.ngll * first.ngll;
  BatchImage out;
  out.n_models = static_cast<int>(models.size());
  auto img = std::make_unique<SetupImage>();
  img->ngll = first.ngll;
  img->axisym = first.axisym;
  img->azimuthal_order = first.azimuthal_order;
  img->D = first.D;
  img->Dj = first.Dj;
  img->w2 = first.w2;
  img->time = first.time;
  img->time.energy_every = 0;
  const std::size_t nd = first.domains.size();
  out.point_model.assign(nd, {});
  // point and element offsets of each model in each union domain
  std::vector<std::vector<int32_t>> goff(nd, std::vector<int32_t>(models.size())),
      eoff(nd, std::vector<int32_t>(models.size()));

  for (std::size_t di = 0; di < nd; ++di) {
    SetupImage::Domain U;
    const SetupImage::Domain& d0 = first.domains[di];
    U.medium = d0.medium;
    U.ncomp = d0.ncomp;
    U.azimuthal_order = d0.azimuthal_order;
    U.axis_components = d0.axis_components;
    DofMap& D = U.dofs;
    D.ngll = first.ngll;
    D.npe = npe;
    D.t_offsets.assign(1, 0);
    bool weighted = false;
    for (const SetupImage* m : models) weighted |= !m->domains[di].dofs.t_weights.empty();
    for (std::size_t mi = 0; mi < models.size(); ++mi) {
      const SetupImage::Domain& d = models[mi]->domains[di];
      const DofMap& M = d.dofs;
      const int32_t g0 = D.n_global, e0 = D.n_elements, r0 = D.n_rows();
      const int32_t q0 = static_cast<int32_t>(D.row_dofs.size());
      const int32_t t0 = D.t_offsets.back();
      goff[di][mi] = g0;
      eoff[di][mi] = e0;
      for (int32_t v : M.l2g) D.l2g.push_back(v >= 0 ? v + g0 : v - r0);
      for (std::size_t r = 1; r < M.row_offsets.size(); ++r) D.row_offsets.push_back(M.row_offsets[r] + q0);
      append_shifted(D.row_dofs, M.row_dofs, g0);
      append(D.row_weights, M.row_weights);
      append_shifted(D.primary, M.primary, e0 * npe);
      for (std::size_t g = 1; g < M.t_offsets.size(); ++g) D.t_offsets.push_back(M.t_offsets[g] + t0);
      append_shifted(D.t_slots, M.t_slots, e0 * npe);
      if (weighted) {
        if (M.t_weights.empty())
          D.t_weights.insert(D.t_weights.end(), M.t_slots.size(), 1.0);
        else
          append(D.t_weights, M.t_weights);
      }
      D.n_elements += M.n_elements;
      D.n_global += M.n_global;
      append(U.mass, d.mass);
      append(U.inv_mass, d.inv_mass);
      append_shifted(U.absorbing_points, d.absorbing_points, g0);
      append(U.absorbing_coef, d.absorbing_coef);
      append_shifted(U.dirichlet_points, d.dirichlet_points, g0);
      append_shifted(U.axis_points, d.axis_points, g0);
      out.point_model[di].insert(out.point_model[di].end(), M.n_global, static_cast<int32_t>(mi));
      // coords [2][n_global] of the union are rebuilt below
    }
    // coordinates
    U.coords.resize(2 * static_cast<std::size_t>(D.n_global));
    for (std::size_t mi = 0; mi < models.size(); ++mi) {
      const SetupImage::Domain& d = models[mi]->domains[di];
      const std::size_t ng = d.dofs.n_global;
      for (std::size_t g = 0; g < ng; ++g) {
        U.coords[goff[di][mi] + g] = d.coords[g];
        U.coords[D.n_global + goff[di][mi] + g] = d.coords[ng + g];
      }
    }
    // blocks: one union block per kind present, in the per-model block order
    std::vector<int> kinds;
    for (const SetupImage* m : models)
      for (const auto& b : m->domains[di].blocks) {
        bool seen = false;
        for (int k : kinds) seen |= k == b.kind;
        if (!seen) kinds.push_back(b.kind);
      }
    for (int kind : kinds) {
      SetupImage::Block B;
      B.kind = kind;
      B.azimuthal_order = d0.azimuthal_order;
      for (std::size_t mi = 0; mi < models.size(); ++mi)
        for (const auto& b : models[mi]->domains[di].blocks) {
          if (b.kind != kind) continue;
          append_shifted(B.elements, b.elements, eoff[di][mi]);
          append(B.metric, b.metric);
          append(B.coef0, b.coef0);
          append(B.coef1, b.coef1);
          append(B.wr, b.wr);
          append(B.hoop, b.hoop);
          append(B.r, b.r);
          append(B.axial, b.axial);
        }
      U.blocks.push_back(std::move(B));
    }
    img->domains.push_back(std::move(U));
  }

  // coupling
  const int fi = first.domain_index(Medium::Acoustic), si = first.domain_index(Medium::Elastic);
  for (std::size_t mi = 0; mi < models.size(); ++mi) {
    const auto& c = models[mi]->coupling;
    if (fi >= 0) append_shifted(img->coupling.fluid_points, c.fluid_points, goff[fi][mi]);
    if (si >= 0) append_shifted(img->coupling.solid_points, c.solid_points, goff[si][mi]);
    append(img->coupling.cx, c.cx);
    append(img->coupling.cz, c.cz);
  }
  // sources and receivers
  for (std::size_t mi = 0; mi < models.size(); ++mi) {
    for (const auto& s : models[mi]->sources) {
      SetupImage::SourceData u = s;
      u.points.clear();
      append_shifted(u.points, s.points, goff[s.domain][mi]);
      img->sources.push_back(std::move(u));
      out.source_model.push_back(static_cast<int32_t>(mi));
    }
    out.receiver_offset.push_back(static_cast<int32_t>(img->receivers.size()));
    for (const auto& r : models[mi]->receivers) {
      SetupImage::ReceiverData u = r;
      u.name = "m" + std::to_string(mi) + ":" + r.name;
      u.points.clear();
      append_shifted(u.points, r.points, goff[r.domain][mi]);
      img->receivers.push_back(std::move(u));
      out.receiver_model.push_back(static_cast<int32_t>(mi));
    }
  }
  out.receiver_offset.push_back(static_cast<int32_t>(img->receivers.size()));
  img->refresh_views();
  out.image = std::move(img);
  return out;
}

PowerJob::PowerJob(const SetupImage& model) {
  check_hip(insem_hip_create(&model.view, &ctx_));
  try {
    const std::vector<float> start = model.power_start_vectors();
    check_hip(insem_hip_lambda_start(ctx_, start.data()));
  } catch (...) {
    insem_hip_destroy(ctx_);
    ctx_ = nullptr;
    throw;
  }
}

StagedPower::StagedPower(const SetupImage& model, void* block) {
  check_hip(insem_hip_stage(&model.view, block, &ctx));
  start = model.power_start_vectors();
}

StagedPower::~StagedPower() {
  if (ctx) insem_hip_destroy(ctx);
}

PowerJob::PowerJob(StagedPower& staged) {
  INSEM_REQUIRE(staged.ctx, "the staged power iteration was already started");
  check_hip(insem_hip_realize(staged.ctx));
  ctx_ = staged.ctx;
  staged.ctx = nullptr;
  try {
    check_hip(insem_hip_lambda_start(ctx_, staged.start.data()));
  } catch (...) {
    insem_hip_destroy(ctx_);
    ctx_ = nullptr;
    throw;
  }
}

PowerJob::~PowerJob() {
  if (ctx_) insem_hip_destroy(ctx_);
}

bool PowerJob::ready() const {
  if (!ctx_) return true;
  int32_t idle = 0;
  check_hip(insem_hip_query(ctx_, &idle));
  return idle != 0;
}

double PowerJob::lambda() {
  if (ctx_) {
    insem_hip_ctx* c = ctx_;
    ctx_ = nullptr;
    const int rc = insem_hip_lambda_result(c, &lambda_);
    insem_hip_destroy(c);
    check_hip(rc);
  }
  return lambda_;
}

insem_hip_ctx* PowerJob::take_context(double* lambda) {
  INSEM_REQUIRE(ctx_, "the power iteration's context was already released");
  const int rc = insem_hip_lambda_result(ctx_, &lambda_);
  if (rc != INSEM_HIP_OK) {
    insem_hip_destroy(ctx_);
    ctx_ = nullptr;
    check_hip(rc);
  }
  insem_hip_ctx* c = ctx_;
  ctx_ = nullptr;
  *lambda = lambda_;
  return c;
}

BatchSolver::BatchSolver(std::vector<std::unique_ptr<SetupImage>> models) : models_(std::move(models)) {
  INSEM_REQUIRE(!models_.empty(), "a batch needs at least one model");
  for (const auto& m : models_) INSEM_REQUIRE(m, "null model in a batch");
  // Each model's own power iteration on a short-lived context of it alone,
  // all on the calling thread.  A window of them is queued before the first
  // result is read, so that the iterations of different models overlap on
  // the device while the host builds the next contexts.
  constexpr std::size_t kWindow = 16;
  std::vector<double> lambdas;
  for (std::size_t w0 = 0; w0 < models_.size(); w0 += kWindow) {
    const std::size_t w1 = std::min(models_.size(), w0 + kWindow);
    std::vector<std::unique_ptr<PowerJob>> jobs;
    for (std::size_t i = w0; i < w1; ++i) jobs.push_back(std::make_unique<PowerJob>(*models_[i]));
    for (auto& j : jobs) lambdas.push_back(j->lambda());
  }
  start(lambdas);
}

BatchSolver::BatchSolver(std::vector<std::unique_ptr<SetupImage>> models, const std::vector<double>& lambdas)
    : models_(std::move(models)) {
  INSEM_REQUIRE(!models_.empty(), "a batch needs at least one model");
  for (const auto& m : models_) INSEM_REQUIRE(m, "null model in a batch");
  INSEM_REQUIRE(lambdas.size() == models_.size(),
                "a batch of " << models_.size() << " models needs as many lambda_max values, not " << lambdas.size());
  start(lambdas);
}

void BatchSolver::start(const std::vector<double>& lambdas) {
  const std::size_t n = models_.size();
  std::vector<ModelTime> own;
  for (std::size_t i = 0; i < n; ++i) {
    own.push_back(model_time(models_[i]->time, lambdas[i]));
    max_steps_ = std::max(max_steps_, own.back().n_steps);
  }
  // union order: longest first (ties: the caller's order)
  std::vector<int> order(n);
  for (std::size_t i = 0; i < n; ++i) order[i] = static_cast<int>(i);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return own[a].n_steps > own[b].n_steps; });
  std::vector<std::unique_ptr<SetupImage>> sorted;
  slot_.assign(n, 0);
  for (std::size_t u = 0; u < n; ++u) {
    sorted.push_back(std::move(models_[order[u]]));
    time_.push_back(own[order[u]]);
    slot_[order[u]] = static_cast<int>(u);
  }
  models_ = std::move(sorted);
  std::vector<const SetupImage*> ptrs;
  for (const auto& m : models_) ptrs.push_back(m.get());
  union_ = union_of(ptrs);
  const SetupImage& U = *union_.image;
  // each model's source rows at its own dt, padded to the union's length
  const std::size_t len = static_cast<std::size_t>(max_steps_) + 1;
  std::vector<float> amps(U.sources.size() * len, 0.0f);
  std::size_t row = 0;
  std::vector<double> dts;
  std::vector<int32_t> steps;
  for (std::size_t mi = 0; mi < models_.size(); ++mi) {
    const std::vector<float> a = models_[mi]->source_amplitudes(time_[mi].dt, time_[mi].n_steps);
    const std::size_t n = static_cast<std::size_t>(time_[mi].n_steps) + 1;
    for (std::size_t s = 0; s < models_[mi]->sources.size(); ++s, ++row)
      std::copy(a.begin() + s * n, a.begin() + (s + 1) * n, amps.begin() + row * len);
    dts.push_back(time_[mi].dt);
    steps.push_back(time_[mi].n_steps);
  }
  insem_hip_batch b{};
  b.n_models = n_models();
  for (std::size_t di = 0; di < union_.point_model.size() && di < 2; ++di) b.point_model[di] = union_.point_model[di].data();
  b.receiver_model = union_.receiver_model.data();
  check_hip(insem_hip_create_batch(&U.view, &b, &ctx_));
  try {
    check_hip(insem_hip_set_time_batch(ctx_, dts.data(), steps.data(), amps.empty() ? nullptr : amps.data()));
  } catch (...) {
    insem_hip_destroy(ctx_);
    ctx_ = nullptr;
    throw;
  }
}

BatchSolver::~BatchSolver() {
  if (ctx_) insem_hip_destroy(ctx_);
}

int BatchSolver::step() const {
  int32_t s = 0;
  check_hip(insem_hip_step(ctx_, &s));
  return s;
}

void BatchSolver::advance(int n) { check_hip(insem_hip_advance(ctx_, n)); }

int BatchSolver::n_samples(int m) const {
  int32_t ns = 0;
  check_hip(insem_hip_n_samples(ctx_, &ns));
  if (ns == 0) return 0;
  const int own = n_steps(m) / union_.image->time.record_every + 1;
  return std::min(static_cast<int>(ns), own);
}

// Model m's traces from the union's trace rows [0, ns) (all receivers'
// quantities, as insem_hip_read_traces lays them out).
std::vector<Solver::Trace> BatchSolver::model_traces(const std::vector<double>& all, int ns, int m) const {
  const SetupImage& U = *union_.image;
  const int keep = std::min(ns, n_steps(m) / U.time.record_every + 1);
  std::vector<Solver::Trace> out;
  std::size_t o = 0;
  const int u = slot(m);
  const int r0 = union_.receiver_offset.at(u), r1 = union_.receiver_offset.at(u + 1);
  for (int ri = 0; ri < static_cast<int>(U.receivers.size()); ++ri) {
    const auto& r = U.receivers[ri];
    const auto& own = models_[u]->receivers;
    for (std::size_t q = 0; q < r.quantities.size(); ++q) {
      const std::size_t n = static_cast<std::size_t>(r.ncomp[q]) * ns;
      if (ri >= r0 && ri < r1) {
        Solver::Trace t;
        t.receiver = own[ri - r0].name;
        t.quantity = r.quantities[q];
        t.ncomp = r.ncomp[q];
        t.data.assign(all.begin() + o, all.begin() + o + static_cast<std::size_t>(t.ncomp) * keep);
        out.push_back(std::move(t));
      }
      o += n;
    }
  }
  return out;
}

std::size_t BatchSolver::trace_values(int ns) const {
  std::size_t total = 0;
  for (const auto& r : union_.image->receivers)
    for (int nc : r.ncomp) total += static_cast<std::size_t>(nc) * ns;
  return total;
}

std::vector<Solver::Trace> BatchSolver::traces(int m) const {
  int32_t ns = 0;
  check_hip(insem_hip_n_samples(ctx_, &ns));
  std::vector<double> all(trace_values(ns) + 1);
  check_hip(insem_hip_read_traces(ctx_, all.data()));
  return model_traces(all, ns, m);
}

int BatchSolver::device_step() const {
  int32_t s = 0;
  check_hip(insem_hip_progress(ctx_, &s));
  return s;
}

std::vector<Solver::Trace> BatchSolver::traces_done(int m) const {
  int32_t ns = 0;
  check_hip(insem_hip_n_samples(ctx_, &ns));  // at least the finished rows
  std::vector<double> all(trace_values(ns) + 1);
  int32_t done = 0;
  check_hip(insem_hip_read_traces_done(ctx_, all.data(), &done));
  INSEM_REQUIRE(done >= n_steps(m) / union_.image->time.record_every + 1,
                "model " << m << " of the batch has not finished its steps on the device");
  return model_traces(all, done, m);
}

int64_t BatchSolver::footprint() const {
  int64_t b = 0;
  check_hip(insem_hip_footprint(ctx_, &b));
  return b;
}

}  // namespace insem_amd
