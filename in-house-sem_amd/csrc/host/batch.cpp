#include "batch.hpp"

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

}  // namespace insem_amd
