// Point-wise kernels of the time step on gfx950: deterministic assembly
// (one work-item per global point, contributions summed in ascending slot
// order from zero, no atomics), the inverse mass, Stacey, Dirichlet/axis and
// coupling treatment, the central-difference corrector and predictor,
// receivers, energy partials and the power-iteration pieces.
//
// Reference: cpp/include/insem/ops/dof_map.hpp:122-141 (assembly),
// cpp/src/solver/solver.cpp:669-762 (sources, coupling, solve),
// 917-963 (predictor, corrector, particle velocity, record), 900-915
// (kinetic energy), 583-667 (power iteration).  Float expressions keep the
// reference's evaluation order; FP64 appears only in the energy and power-
// iteration reductions.

#include "../../../include/insem_hip.h"
#include "device.hpp"
#include "launch.hpp"
#include "receivers.hpp"

namespace insem_hip {
namespace {

constexpr int kTB = 256;

// Fixed-tree block sum of one double per work-item; result in thread 0.
__device__ double block_sum(double v, double* lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if (t < w) lds[t] += lds[t + w];
    __syncthreads();
  }
  const double r = lds[0];
  __syncthreads();
  return r;
}

// Assembled force of point g from f_local through its CSR row, summed in
// slot order.  The row is read in chunks of kChunk slots whose addresses,
// then values, are all loaded before the chunk's adds.
template <int C, bool W>
__device__ __forceinline__ void gather_rows(const DomDev& d, int32_t s0, int32_t s1, float* f) {
  constexpr int kChunk = 4;
#pragma unroll
  for (int c = 0; c < C; ++c) f[c] = 0.0f;
  for (int32_t s = s0; s < s1; s += kChunk) {
    uint32_t at[kChunk];
    float w[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const bool in = s + q < s1;
      at[q] = in ? *at32(d.t_addr, s + q) : 0;
      if constexpr (W) w[q] = in ? *at32(d.t_w, s + q) : 0.0f;
    }
    float p[kChunk][C];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
#pragma unroll
      for (int c = 0; c < C; ++c) p[q][c] = s + q < s1 ? *at32(d.fl, at[q] + c * d.npe) : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (s + q < s1) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
          if constexpr (W)
            f[c] += w[q] * p[q][c];
          else
            f[c] += p[q][c];
        }
      }
    }
  }
}

template <int C, bool W>
__device__ __forceinline__ void gather_sum(const DomDev& d, uint32_t g, float* f) {
  gather_rows<C, W>(d, *at32(d.t_off, g), *at32(d.t_off, g + 1), f);
}

// Coefficients of point g at `step`: the launch's, or on a batch its
// model's.  A batch model past its last step is inactive, and its last step
// predicts nothing (as a run of that length alone).
struct PointStep {
  float rdt, hdt, hdt2;
  double cc;
  int pred;
  bool active;
};

__device__ __forceinline__ PointStep point_step(const DomDev& d, const StepCoef& sc, int pred, int step, int64_t g,
                                                int bm) {
  PointStep p{sc.rdt, sc.hdt, sc.hdt2, sc.cc, pred, true};
  if (d.models) {
    auto take = [&](const ModelCoef& m) {
      p.rdt = m.rdt;
      p.hdt = m.hdt;
      p.hdt2 = m.hdt2;
      p.cc = m.cc;
      p.active = step <= m.n_steps;
      p.pred = pred && step < m.n_steps;
    };
    if (bm >= 0)
      take(d.models[bm]);  // uniform
    else
      take(d.models[*at32(d.pmodel, static_cast<uint32_t>(g))]);
  }
  return p;
}

// On a batch: the one model of the points of the tail block [g0, g0 + kTB)
// when they share one (the usual case), else -1; uniform over the block.
// (Points of models past their last step do nothing: PointStep::active.)
__device__ __forceinline__ int tail_block_model(const DomDev& d, int64_t g0) {
  if (!d.chunk_model) return -1;
  const int64_t g1 = (g0 + kTB < d.ng ? g0 + kTB : d.ng) - 1;
  if (g1 < g0) return -1;
  const int32_t m0 = d.chunk_model[g0 >> 8], m1 = d.chunk_model[g1 >> 8];
  return m0 == m1 ? m0 : -1;
}

// The values a regular point's update reads besides its force, loaded
// before the force is assembled so that both sets of loads are in flight
// together.
template <int C>
struct PointState {
  float im;
  float u[C], v[C];
};

template <int C>
__device__ __forceinline__ PointState<C> load_point(const DomDev& d, int64_t g) {
  const uint32_t ng = static_cast<uint32_t>(d.ng), gi = static_cast<uint32_t>(g);
  PointState<C> s;
  s.im = *at32(d.im, gi);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    s.u[c] = *at32(d.u, c * ng + gi);
    s.v[c] = *at32(d.v, c * ng + gi);
  }
  return s;
}

// a = M^-1 f, optional corrector, energy terms and the next predictor at a
// regular point g with assembled force f and the values st loaded before.
template <int C>
__device__ __forceinline__ void update_point(const DomDev& d, const PointStep& ps, int corr, bool energy, int64_t g,
                                             const float* f, const PointState<C>& st, double& ep, double& ek) {
  const uint32_t ng = static_cast<uint32_t>(d.ng), gi = static_cast<uint32_t>(g);
  const float im = st.im;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint32_t o = c * ng + gi;
    float u = st.u[c], v = st.v[c];
    const float a = f[c] * im;
    if (energy) ep += static_cast<double>(u) * static_cast<double>(f[c]);
    if (corr) v = v + ps.hdt * a;
    if (energy) {
      const double vd = v, ad = a;
      ek += static_cast<double>(*at32(d.mass, gi)) * (0.5 * vd * vd - ps.cc * ad * ad);
    }
    if (ps.pred) {
      // The next step recomputes a at regular points before anything
      // reads it, so it is only stored when the step ends the call.
      u = u + (ps.rdt * v + ps.hdt2 * a);
      v = v + ps.hdt * a;
      *at32(d.u, o) = u;
    } else {
      *at32(d.a, o) = a;
    }
    *at32(d.v, o) = v;
  }
}

__device__ __forceinline__ void write_partials(bool energy, double ep, double ek, double* lds, double* partial) {
  if (!energy) return;
  const double sp = block_sum(ep, lds);
  const double sk = block_sum(ek, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sp;
    partial[2 * blockIdx.x + 1] = sk;
  }
}

// Regular points [first, n_reg) assembled through the transpose CSR; vb is
// the block's index within this part of the launch.
template <int C, bool W>
__device__ __forceinline__ void regular_body(const DomDev& d, int64_t first, const StepCoef& sc, int corr, int pred,
                                             int step, bool energy, int vb, double& ep, double& ek) {
  const int64_t g0 = first + static_cast<int64_t>(vb) * kTB, g = g0 + threadIdx.x;
  if (g >= d.n_reg) return;
  const PointStep ps = point_step(d, sc, pred, step, g, tail_block_model(d, g0));
  if (!ps.active) return;
  const PointState<C> st = load_point<C>(d, g);
  float f[C];
  gather_sum<C, W>(d, g, f);
  update_point<C>(d, ps, corr, energy, g, f, st, ep, ek);
}

// Sums of the K contributions per component of a grouped point, from zero
// in rank order: rank r of component c in row c * rows + r (stride n); all
// C * K loads are issued before the first add.
template <int C, int K>
__device__ __forceinline__ void group_sum(const float* contrib, uint32_t off, uint32_t n, uint32_t rows, float* f) {
  float x[C * K];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int r = 0; r < K; ++r) x[c * K + r] = *at32(contrib, off + (c * rows + r) * n);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < K; ++r) s += x[c * K + r];
    f[c] = s;
  }
}

// Grouped points: the contributions of a point lie in rank order in the
// group rows; the sum starts from zero in slot order as the assembly's does.
template <int C>
__device__ __forceinline__ void grouped_body(const DomDev& d, const StepCoef& sc, int corr, int pred, int step,
                                             bool energy, int vb, double& ep, double& ek) {
  const int64_t g0 = d.grp.begin + static_cast<int64_t>(vb) * kTB, g = g0 + threadIdx.x;
  if (g >= d.grp.end) return;
  const PointStep ps = point_step(d, sc, pred, step, g, tail_block_model(d, g0));
  if (!ps.active) return;
  const uint32_t o = static_cast<uint32_t>(g - d.grp.begin);
  const uint32_t z = static_cast<uint32_t>(d.grp.end - d.grp.begin), rows = static_cast<uint32_t>(d.grp.rows);
  const PointState<C> st = load_point<C>(d, g);
  float f[C];
  switch (group_count(d.grp, static_cast<uint32_t>(g))) {
    case 2: group_sum<C, 2>(d.contrib, o, z, rows, f); break;
    case 3: group_sum<C, 3>(d.contrib, o, z, rows, f); break;
    case 4: group_sum<C, 4>(d.contrib, o, z, rows, f); break;
    case 5: group_sum<C, 5>(d.contrib, o, z, rows, f); break;
    case 6: group_sum<C, 6>(d.contrib, o, z, rows, f); break;
    case 7: group_sum<C, 7>(d.contrib, o, z, rows, f); break;
    default: group_sum<C, 8>(d.contrib, o, z, rows, f); break;
  }
  update_point<C>(d, ps, corr, energy, g, f, st, ep, ek);
}

// Small SPD solve without pivoting, in double as the reference's
// (solver.cpp:67-80); it runs at Stacey points only.
template <int N>
__device__ __forceinline__ void solve_small(double* A, double* b) {
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double l = A[i * N + k] / A[k * N + k];
#pragma unroll
      for (int j = k; j < N; ++j) A[i * N + j] -= l * A[k * N + j];
      b[i] -= l * b[k];
    }
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    double s = b[k];
#pragma unroll
    for (int j = k + 1; j < N; ++j) s -= A[k * N + j] * b[j];
    b[k] = s / A[k * N + k];
  }
}

// Special points: sources, coupling, Stacey, Dirichlet and axis conditions;
// the pre-predictor values go to the side buffers.
// Coupling operand of a special point (the other domain's values at the
// partner point) when the caller has it; else it is read from sp.other.
struct Operand {
  bool given = false;
  float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
};
// The point's new acceleration (components beyond C are 0); active is
// false when the point's model is past its last step.
struct Accel {
  bool active = false;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
};

// Everything a special point's update reads, loaded before any of its
// arithmetic.
template <int C>
struct SpecialIn {
  uint32_t g = 0;
  int32_t q0 = 0, q1 = 0, ci = -1, ai = -1;
  uint8_t fl = 0;
  float u[C], v[C];
  float im = 0.0f;
  double mass = 0.0;
  PointStep ps;
  double h_damp = 0.0;
  float cx = 0.0f, cz = 0.0f;
  uint32_t o = 0;  // the coupling partner's point in the other domain
  float cf[C * C];
  float ot[3] = {0.0f, 0.0f, 0.0f};
  float f[C];
};

// The loads of special point i, issued level by level (first everything
// indexed by the point, then what those indices select; coupling operands
// with a clamped index), so that each wait covers one level of the chains.
// `other`: read the coupling operand from the other domain (else the caller
// supplies it to special_finish).
template <int C, bool W>
__device__ __forceinline__ void special_load(const DomDev& d, const SpecialDev& sp, const StepCoef& sc, int pred,
                                             int step, int64_t i, int bm, bool other, SpecialIn<C>& in) {
  const uint32_t g = static_cast<uint32_t>(d.n_reg + i), ng = static_cast<uint32_t>(d.ng);
  in.g = g;
  const int32_t t0 = *at32(d.t_off, g), t1 = *at32(d.t_off, g + 1);
  in.q0 = sp.src_off[i];
  in.q1 = sp.src_off[i + 1];
  in.ci = sp.cpl_idx[i];
  in.ai = sp.abs_idx[i];
  in.fl = sp.flags[i];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    in.u[c] = *at32(d.u, c * ng + g);
    in.v[c] = *at32(d.v, c * ng + g);
  }
  in.im = *at32(d.im, g);
  in.mass = *at32(d.mass, g);
  in.ps = point_step(d, sc, pred, step, g, bm);
  if (sp.cpl_cx) {
    const int32_t cq = in.ci >= 0 ? in.ci : 0;
    in.cx = sp.cpl_cx[cq];
    in.cz = sp.cpl_cz[cq];
    in.o = static_cast<uint32_t>(sp.cpl_other[cq]);
  }
  if (in.ai >= 0) {
#pragma unroll
    for (int r = 0; r < C * C; ++r) in.cf[r] = sp.abs_coef[static_cast<int64_t>(in.ai) * C * C + r];
  }
  if (other && in.ci >= 0) {
    const uint32_t on = static_cast<uint32_t>(sp.other_ng);
    in.ot[0] = *at32(sp.other, in.o);
    if (sp.cpl_role != 1) {
      in.ot[1] = *at32(sp.other, on + in.o);
      if (sp.pm) in.ot[2] = *at32(sp.other, 2 * on + in.o);
    }
  }
  in.h_damp = sc.h_damp;
  if (d.models && sc.h_damp > 0.0) {
    if (bm >= 0)
      in.h_damp = d.models[bm].h_damp;  // uniform
    else
      in.h_damp = d.models[*at32(d.pmodel, g)].h_damp;
  }
  gather_rows<C, W>(d, t0, t1, in.f);
}

// The update of a special point from its loads: sources, coupling, Stacey,
// Dirichlet and axis conditions; the pre-predictor values go to the side
// buffers.  ot_in (when given) is the coupling operand.
template <int C, bool W>
__device__ __forceinline__ Accel special_finish(const DomDev& d, const SpecialDev& sp, const StepCoef& sc, int corr,
                                                int step, bool energy, int64_t i, SpecialIn<C>& in, double& ep,
                                                double& ek, Operand ot_in) {
  if (!in.ps.active) return Accel{};
  const uint32_t g = in.g, ng = static_cast<uint32_t>(d.ng);
  const PointStep& ps = in.ps;
  float* u = in.u;
  float* v = in.v;
  float* f = in.f;
  float a[C];
  if (ot_in.given) {
    in.ot[0] = ot_in.v0;
    in.ot[1] = ot_in.v1;
    in.ot[2] = ot_in.v2;
  }
  const float* ot = in.ot;
  const float cx = in.cx, cz = in.cz;
  const int32_t ci = in.ci, ai = in.ai;
  const uint8_t fl = in.fl;
  const float im = in.im;
  const double mass = in.mass;
  const float* cf = in.cf;
  const double h_damp = in.h_damp;
#pragma unroll
  for (int c = 0; c < C; ++c)
    if (energy) ep += static_cast<double>(u[c]) * static_cast<double>(f[c]);
  // Sources, in source order (solver.cpp:679-687).
  for (int32_t q = in.q0; q < in.q1; ++q) {
    const float amp = sp.amps[static_cast<int64_t>(sp.src_id[q]) * sp.amp_stride + step];
#pragma unroll
    for (int c = 0; c < C; ++c) f[c] += sp.src_coef[static_cast<int64_t>(q) * C + c] * amp;
  }
  // Coupling (solver.cpp:693-727).
  if (ci >= 0) {
    const float h = 0.70710678118654752440f;
    if (sp.cpl_role == 1) {
      if constexpr (C >= 2) {
        const float p = ot[0];
        if constexpr (C == 3) {
          f[0] += h * cx * p;
          f[1] += h * cx * p;
        } else {
          f[0] += cx * p;
        }
        f[C - 1] += cz * p;
      }
    } else {
      const float ar = sp.pm ? h * (ot[0] + ot[1]) : ot[0];
      const float az = sp.pm ? ot[2] : ot[1];
      f[0] -= cx * ar + cz * az;
    }
  }
  // a = M^-1 f, Stacey points by (M + h C) a = f - C v (solver.cpp:729-762).
#pragma unroll
  for (int c = 0; c < C; ++c) a[c] = f[c] * im;
  if (ai >= 0) {
    double A[C * C], rhs[C];
#pragma unroll
    for (int r = 0; r < C; ++r) {
      rhs[r] = f[r];
#pragma unroll
      for (int s = 0; s < C; ++s) {
        rhs[r] -= static_cast<double>(cf[r * C + s]) * v[s];
        A[r * C + s] = h_damp * cf[r * C + s] + (r == s ? mass : 0.0);
      }
    }
    solve_small<C>(A, rhs);
#pragma unroll
    for (int c = 0; c < C; ++c) a[c] = static_cast<float>(rhs[c]);
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    if ((fl & 1) || ((fl & 2) && ((sp.axis_mask >> c) & 1))) a[c] = 0.0f;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    if (corr) v[c] = v[c] + ps.hdt * a[c];
    if (energy) {
      const double vd = v[c], ad = a[c];
      ek += mass * (0.5 * vd * vd - ps.cc * ad * ad);
    }
    sp.su[c * sp.n_sp + i] = u[c];
    sp.sv[c * sp.n_sp + i] = v[c];
    sp.sa[c * sp.n_sp + i] = a[c];
    const uint32_t oo = c * ng + g;
    if (ps.pred) {
      *at32(d.u, oo) = u[c] + (ps.rdt * v[c] + ps.hdt2 * a[c]);
      v[c] = v[c] + ps.hdt * a[c];
    }
    *at32(d.v, oo) = v[c];
    *at32(d.a, oo) = a[c];
  }
  Accel r;
  r.active = true;
  r.a0 = a[0];
  if constexpr (C >= 2) r.a1 = a[1];
  if constexpr (C >= 3) r.a2 = a[2];
  return r;
}

// A special point's update (loads, then arithmetic).
template <int C, bool W>
__device__ __forceinline__ Accel special_point(const DomDev& d, const SpecialDev& sp, const StepCoef& sc, int corr,
                                               int pred, int step, bool energy, int64_t i, int bm, double& ep,
                                               double& ek, Operand ot_in) {
  SpecialIn<C> in;
  special_load<C, W>(d, sp, sc, pred, step, i, bm, !ot_in.given, in);
  return special_finish<C, W>(d, sp, sc, corr, step, energy, i, in, ep, ek, ot_in);
}

template <int C, bool W>
__device__ __forceinline__ void special_body(const DomDev& d, const SpecialDev& sp, const StepCoef& sc, int corr,
                                             int pred, int step, bool energy, int vb, double& ep, double& ek) {
  const int64_t i = static_cast<int64_t>(vb) * kTB + threadIdx.x;
  if (i >= sp.n_sp) return;
  const int bm = tail_block_model(d, d.n_reg + static_cast<int64_t>(vb) * kTB);
  special_point<C, W>(d, sp, sc, corr, pred, step, energy, i, bm, ep, ek, Operand{});
}


// Coupled model, solid special point i: the solid point, then, when it is on
// the interface, its fluid partner with the solid's new acceleration as the
// coupling operand (the reference's order: the fluid's coupling term reads
// the solid's new a, solver.cpp:669-727).  The fluid partner is updated by
// this work-item only, so the solid's read of its pressure precedes its
// update.
template <int CS, bool WS, int CF, bool WF>
__device__ __forceinline__ void special_pair_body(const DomDev& ds, const SpecialDev& sps, const DomDev& df,
                                                  const SpecialDev& spf, const StepCoef& sc, int corr, int pred,
                                                  int step, bool energy, int vb, double& ep, double& ek) {
  const int64_t i = static_cast<int64_t>(vb) * kTB + threadIdx.x;
  if (i >= sps.n_sp) return;
  // (a fluid partner belongs to its solid point's model)
  const int bm = tail_block_model(ds, ds.n_reg + static_cast<int64_t>(vb) * kTB);
  SpecialIn<CS> a;
  special_load<CS, WS>(ds, sps, sc, pred, step, i, bm, true, a);
  const bool partner = a.ci >= 0;
  const int64_t i_f = static_cast<int64_t>(a.o) - df.n_reg;
  SpecialIn<CF> b;
  // With three solid components the partner's loads go out before the solid
  // point's arithmetic (only its coupling term waits for the solid's new
  // acceleration); the tail then needs more registers and runs at five
  // waves per SIMD, which measured faster than six with the two points one
  // after the other.  With two components, or a weighted fluid gather, both
  // points' values at once would spill registers, so the partner loads
  // after the solid point's update.
  constexpr bool overlap = CS == 3 && !WF;
  if constexpr (overlap) {
    if (partner) special_load<CF, WF>(df, spf, sc, pred, step, i_f, bm, false, b);
  }
  const Accel as = special_finish<CS, WS>(ds, sps, sc, corr, step, energy, i, a, ep, ek, Operand{});
  if (!as.active || !partner) return;
  if constexpr (!overlap) special_load<CF, WF>(df, spf, sc, pred, step, i_f, bm, false, b);
  Operand op;
  op.given = true;
  op.v0 = as.a0;
  op.v1 = as.a1;
  op.v2 = as.a2;
  special_finish<CF, WF>(df, spf, sc, corr, step, energy, i_f, b, ep, ek, op);
}

// Coupled model, fluid special points off the interface (the interface ones
// go with their solid partner).
template <int C, bool W>
__device__ __forceinline__ void special_uncoupled_body(const DomDev& d, const SpecialDev& sp, const StepCoef& sc,
                                                       int corr, int pred, int step, bool energy, int vb, double& ep,
                                                       double& ek) {
  const int64_t i = static_cast<int64_t>(vb) * kTB + threadIdx.x;
  if (i >= sp.n_sp || sp.cpl_idx[i] >= 0) return;
  const int bm = tail_block_model(d, d.n_reg + static_cast<int64_t>(vb) * kTB);
  special_point<C, W>(d, sp, sc, corr, pred, step, energy, i, bm, ep, ek, Operand{});
}

// The tail of a step on one domain, one launch: blocks [0, n_sp_blocks) take
// the special points (first, so that their dependent loads start early),
// the next n_grp_blocks the grouped points, the rest the regular points
// [reg_first, n_reg) through the CSR.  Energy partials: [block][2].
// Waves per SIMD of the tail (the kernel is bound by the latency of its
// gathers): all eight the hardware holds for one component; seven for two,
// where eight would spill scalar registers.
constexpr int tail_waves(int c) { return c <= 1 ? 8 : c == 2 ? 7 : 5; }

template <int C, bool W>
__global__ __launch_bounds__(kTB, tail_waves(C)) void k_tail(DomDev d, SpecialDev sp, StepCoef sc, int corr, int pred,
                                                             StepClock clk, double* partial, int n_sp_blocks,
                                                             int n_grp_blocks, int64_t reg_first) {
  __shared__ double lds[kTB];
  const int step = clock_step(clk);
  const bool energy = clock_energy(clk, step);
  double ep = 0.0, ek = 0.0;
  const int b = blockIdx.x;
  if (b < n_sp_blocks)
    special_body<C, W>(d, sp, sc, corr, pred, step, energy, b, ep, ek);
  else if (b < n_sp_blocks + n_grp_blocks)
    grouped_body<C>(d, sc, corr, pred, step, energy, b - n_sp_blocks, ep, ek);
  else
    regular_body<C, W>(d, reg_first, sc, corr, pred, step, energy, b - n_sp_blocks - n_grp_blocks, ep, ek);
  write_partials(energy, ep, ek, lds, partial);
}

// The tails of both domains of a coupled model in one launch: blocks
// [0, nss) the solid's special points (with their fluid partners on the
// interface), [nss, nss + nsf) the fluid's other special points, then the
// grouped and regular points of the solid and of the fluid.
template <int CS, bool WS, int CF, bool WF>
__global__ __launch_bounds__(kTB, tail_waves(CS)) void k_tail_pair(DomDev ds, SpecialDev sps, DomDev df,
                                                                   SpecialDev spf, StepCoef sc, int corr, int pred,
                                                                   StepClock clk, double* partial, int nss, int nsf,
                                                                   int ngs, int nrs, int ngf) {
  __shared__ double lds[kTB];
  const int step = clock_step(clk);
  const bool energy = clock_energy(clk, step);
  double ep = 0.0, ek = 0.0;
  int b = blockIdx.x;
  if (b < nss) {
    special_pair_body<CS, WS, CF, WF>(ds, sps, df, spf, sc, corr, pred, step, energy, b, ep, ek);
  } else if ((b -= nss) < nsf) {
    special_uncoupled_body<CF, WF>(df, spf, sc, corr, pred, step, energy, b, ep, ek);
  } else if ((b -= nsf) < ngs) {
    grouped_body<CS>(ds, sc, corr, pred, step, energy, b, ep, ek);
  } else if ((b -= ngs) < nrs) {
    regular_body<CS, WS>(ds, ds.grp.end, sc, corr, pred, step, energy, b, ep, ek);
  } else if ((b -= nrs) < ngf) {
    grouped_body<CF>(df, sc, corr, pred, step, energy, b, ep, ek);
  } else {
    regular_body<CF, WF>(df, df.grp.end, sc, corr, pred, step, energy, b - ngf, ep, ek);
  }
  write_partials(energy, ep, ek, lds, partial);
}

// Predictor of the step after the clock's (start of an advance call); on a
// batch with each point's model coefficients, skipping finished models.
__global__ __launch_bounds__(kTB) void k_predictor(DomDev d, StepCoef sc, StepClock clk) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i >= d.ncomp * d.ng) return;
  const PointStep ps = point_step(d, sc, 1, clock_step(clk) + 1, i % d.ng, -1);
  if (!ps.active) return;
  const float vi = d.v[i], ai = d.a[i];
  d.u[i] = d.u[i] + (ps.rdt * vi + ps.hdt2 * ai);
  d.v[i] = vi + ps.hdt * ai;
}

// The same predictor over every value of up to two domains of one model
// (no per-point coefficients): one work-item per four consecutive values of
// a domain's u, v and a (component-major, 16-byte aligned), so the launch
// streams the three fields at full width.  Each value gets the operations
// of k_predictor in the same order.
struct PredSpan {
  float* u = nullptr;
  float* v = nullptr;
  const float* a = nullptr;
  int64_t n = 0;  // values
};

__global__ __launch_bounds__(kTB) void k_predictor_flat(PredSpan s0, PredSpan s1, int64_t q0, int64_t q1, float rdt,
                                                       float hdt, float hdt2) {
  int64_t q = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (q >= q0 + q1) return;
  const bool second = q >= q0;
  const PredSpan& sp = second ? s1 : s0;
  if (second) q -= q0;
  const int64_t i = 4 * q;
  if (i + 4 <= sp.n) {
    const float4 u = *reinterpret_cast<const float4*>(sp.u + i);
    const float4 v = *reinterpret_cast<const float4*>(sp.v + i);
    const float4 a = *reinterpret_cast<const float4*>(sp.a + i);
    float4 un, vn;
    un.x = u.x + (rdt * v.x + hdt2 * a.x);
    un.y = u.y + (rdt * v.y + hdt2 * a.y);
    un.z = u.z + (rdt * v.z + hdt2 * a.z);
    un.w = u.w + (rdt * v.w + hdt2 * a.w);
    vn.x = v.x + hdt * a.x;
    vn.y = v.y + hdt * a.y;
    vn.z = v.z + hdt * a.z;
    vn.w = v.w + hdt * a.w;
    *reinterpret_cast<float4*>(sp.u + i) = un;
    *reinterpret_cast<float4*>(sp.v + i) = vn;
  } else {
    for (int64_t k = i; k < sp.n; ++k) {
      const float vi = sp.v[k], ai = sp.a[k];
      sp.u[k] = sp.u[k] + (rdt * vi + hdt2 * ai);
      sp.v[k] = vi + hdt * ai;
    }
  }
}

template <int C, bool W>
__global__ __launch_bounds__(kTB) void k_assemble(DomDev d) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (g >= d.ng) return;
  float f[C];
  gather_sum<C, W>(d, g, f);
#pragma unroll
  for (int c = 0; c < C; ++c) d.f[c * d.ng + g] = f[c];
}

// One work-group sums n rows of `width` doubles in a fixed order.
__global__ __launch_bounds__(kTB) void k_reduce(const double* partial, int n, int width, double* out, int accumulate) {
  __shared__ double lds[kTB];
  for (int w = 0; w < width; ++w) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kTB) s += partial[static_cast<int64_t>(i) * width + w];
    const double t = block_sum(s, lds);
    if (threadIdx.x == 0) out[w] = accumulate ? out[w] + t : t;
  }
}

// Energy row of an energy step: rows[step / energy_every] = sums of the
// [n][2] partials (u.f, kinetic).
__global__ __launch_bounds__(kTB) void k_energy_row(const double* partial, int n, StepClock clk, double* rows) {
  __shared__ double lds[kTB];
  const int step = clock_step(clk);
  if (!clock_energy(clk, step)) return;
  for (int w = 0; w < 2; ++w) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kTB) s += partial[2 * static_cast<int64_t>(i) + w];
    const double t = block_sum(s, lds);
    if (threadIdx.x == 0) rows[2 * (step / clk.energy_every) + w] = t;
  }
}

__global__ void k_clock_add(int32_t* base, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *base += n;
}

constexpr int kRecWG = 64;

__global__ __launch_bounds__(kRecWG) void k_receivers(ReceiverDev r, StepClock clk, int pv_mode, double half_dt) {
  __shared__ double sw[kRecWG];
  __shared__ float sx[kRecWG];
  receiver_task<kRecWG>(r, blockIdx.x, clk, clock_step(clk), pv_mode, half_dt, sw, sx);
}

__global__ void k_zero_points(float* x, int64_t ng, const int32_t* pts, int n, int comp_mask, int ncomp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncomp; ++c)
    if ((comp_mask >> c) & 1) x[c * ng + pts[i]] = 0.0f;
}

// Power-iteration statistics of one domain (solver.cpp:593-612): u.f,
// sum m x^2, y = -f/m -> x, sum y^2.
__global__ __launch_bounds__(kTB) void k_power_stats(DomDev d, float* x, double* partial, const double* state) {
  __shared__ double lds[kTB];
  if (state[1] != 0.0) return;
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  double uf = 0.0, xmx = 0.0, nrm = 0.0;
  if (g < d.ng) {
    const double m = d.mass[g], im = d.im[g];
    for (int c = 0; c < d.ncomp; ++c) {
      const int64_t o = c * d.ng + g;
      const double xv = x[o], fv = d.f[o];
      uf += xv * fv;
      xmx += m * xv * xv;
      const double y = -fv * im;
      x[o] = static_cast<float>(y);
      nrm += y * y;
    }
  }
  const double s0 = block_sum(uf, lds), s1 = block_sum(xmx, lds), s2 = block_sum(nrm, lds);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = s0;
    partial[3 * blockIdx.x + 1] = s1;
    partial[3 * blockIdx.x + 2] = s2;
  }
}

__global__ void k_power_update(const double* out3, double* state) {
  if (threadIdx.x != 0 || state[1] != 0.0) return;
  const double xkx = -out3[0], xmx = out3[1], nrm = out3[2];
  if (xmx > 0.0 && xkx / xmx > state[0]) state[0] = xkx / xmx;
  if (nrm == 0.0) {
    state[1] = 1.0;
    return;
  }
  state[2] = 1.0 / sqrt(nrm);
}

__global__ __launch_bounds__(kTB) void k_scale(float* x, int64_t n, const double* state) {
  if (state[1] != 0.0) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i < n) x[i] = static_cast<float>(x[i] * state[2]);
}

__global__ __launch_bounds__(kTB) void k_mass_norm(DomDev d, const float* x, double* partial) {
  __shared__ double lds[kTB];
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  double s = 0.0;
  if (g < d.ng) {
    const double m = d.mass[g];
    for (int c = 0; c < d.ncomp; ++c) {
      const double xv = x[c * d.ng + g];
      s += m * xv * xv;
    }
  }
  const double t = block_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kTB) void k_accel(DomDev d) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (g >= d.ng) return;
  const float im = d.im[g];
  for (int c = 0; c < d.ncomp; ++c) d.a[c * d.ng + g] = d.f[c * d.ng + g] * im;
}

__global__ void k_coupling_power(DomDev d, const int32_t* pts, const int32_t* other, const float* cx, const float* cz,
                                 int n, int role, int pm, const float* src, int64_t src_ng) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float h = 0.70710678118654752440f;
  const int64_t g = pts[i], o = other[i];
  if (role == 1) {
    const float p = src[o];
    if (pm) {
      d.f[g] += h * cx[i] * p;
      d.f[d.ng + g] += h * cx[i] * p;
    } else {
      d.f[g] += cx[i] * p;
    }
    d.f[(pm ? 2 : 1) * d.ng + g] += cz[i] * p;
  } else {
    const float ar = pm ? h * (src[o] + src[src_ng + o]) : src[o];
    d.f[g] -= cx[i] * ar + cz[i] * src[(pm ? 2 : 1) * src_ng + o];
  }
}

// Coupled growth (solver.cpp:640-658): nx, ny are mass norms of u and a.
__global__ void k_coupled_update(const double* nx, const double* ny, int it, double* state) {
  if (threadIdx.x != 0 || state[1] != 0.0) return;
  const double a = nx[0], b = ny[0];
  if (a <= 0.0 || b <= 0.0) {
    state[1] = 1.0;
    return;
  }
  if (it >= 60 && sqrt(b / a) > state[0]) state[0] = sqrt(b / a);
  state[2] = -1.0 / sqrt(b);
}

__global__ __launch_bounds__(kTB) void k_scale_from(float* x, const float* a, int64_t n, const double* state) {
  if (state[1] != 0.0) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i < n) x[i] = static_cast<float>(a[i] * state[2]);
}

__global__ __launch_bounds__(kTB) void k_dot(const float* a, const float* b, int64_t n, double* partial) {
  __shared__ double lds[kTB];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  const double v = i < n ? static_cast<double>(a[i]) * static_cast<double>(b[i]) : 0.0;
  const double t = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

inline int blocks_for(int64_t n) { return static_cast<int>((n + kTB - 1) / kTB); }

}  // namespace

int tail_blocks(int64_t n) { return blocks_for(n); }

hipError_t launch_tail(const DomDev& d, const SpecialDev& sp, const StepCoef& sc, int corr, int pred,
                       const StepClock& clk, bool grouped, double* partial, int* n_blocks, hipStream_t s) {
  const int nsp = blocks_for(sp.n_sp);
  const int ngr = grouped ? blocks_for(d.grp.end - d.grp.begin) : 0;
  const int64_t first = grouped ? d.grp.end : 0;
  const int nrg = first < d.n_reg ? blocks_for(d.n_reg - first) : 0;
  const int nb = nsp + ngr + nrg;
  if (n_blocks) *n_blocks = nb;
  if (nb == 0) return hipSuccess;
  const bool w = d.t_w != nullptr;
#define INSEM_TAIL(C)                                                                                          \
  if (w)                                                                                                       \
    hipLaunchKernelGGL((k_tail<C, true>), dim3(nb), dim3(kTB), 0, s, d, sp, sc, corr, pred, clk, partial, nsp, \
                       ngr, first);                                                                            \
  else                                                                                                         \
    hipLaunchKernelGGL((k_tail<C, false>), dim3(nb), dim3(kTB), 0, s, d, sp, sc, corr, pred, clk, partial,     \
                       nsp, ngr, first);
  switch (d.ncomp) {
    case 1: INSEM_TAIL(1) break;
    case 2: INSEM_TAIL(2) break;
    case 3: INSEM_TAIL(3) break;
    default: return hipErrorInvalidValue;
  }
#undef INSEM_TAIL
  return hipGetLastError();
}

namespace {

template <int CS, bool WS, int CF, bool WF>
void tail_pair_one(const DomDev& ds, const SpecialDev& sps, const DomDev& df, const SpecialDev& spf,
                   const StepCoef& sc, int corr, int pred, const StepClock& clk, double* partial, int nss, int nsf,
                   int ngs, int nrs, int ngf, int nrf, hipStream_t s) {
  hipLaunchKernelGGL((k_tail_pair<CS, WS, CF, WF>), dim3(nss + nsf + ngs + nrs + ngf + nrf), dim3(kTB), 0, s, ds,
                     sps, df, spf, sc, corr, pred, clk, partial, nss, nsf, ngs, nrs, ngf);
}

}  // namespace

hipError_t launch_tail_pair(const DomDev& ds, const SpecialDev& sps, const DomDev& df, const SpecialDev& spf,
                            const StepCoef& sc, int corr, int pred, const StepClock& clk, double* partial,
                            int* n_blocks, hipStream_t s) {
  if (df.ncomp != 1 || (ds.ncomp != 2 && ds.ncomp != 3) || sps.cpl_role != 1) return hipErrorInvalidValue;
  const int nss = blocks_for(sps.n_sp), nsf = blocks_for(spf.n_sp);
  const int ngs = blocks_for(ds.grp.end - ds.grp.begin), ngf = blocks_for(df.grp.end - df.grp.begin);
  const int nrs = ds.grp.end < ds.n_reg ? blocks_for(ds.n_reg - ds.grp.end) : 0;
  const int nrf = df.grp.end < df.n_reg ? blocks_for(df.n_reg - df.grp.end) : 0;
  const int nb = nss + nsf + ngs + nrs + ngf + nrf;
  if (n_blocks) *n_blocks = nb;
  if (nb == 0) return hipSuccess;
  const bool ws = ds.t_w != nullptr, wf = df.t_w != nullptr;
#define INSEM_TP(CS, WS, WF) \
  tail_pair_one<CS, WS, 1, WF>(ds, sps, df, spf, sc, corr, pred, clk, partial, nss, nsf, ngs, nrs, ngf, nrf, s)
  if (ds.ncomp == 2) {
    if (ws)
      wf ? INSEM_TP(2, true, true) : INSEM_TP(2, true, false);
    else
      wf ? INSEM_TP(2, false, true) : INSEM_TP(2, false, false);
  } else {
    if (ws)
      wf ? INSEM_TP(3, true, true) : INSEM_TP(3, true, false);
    else
      wf ? INSEM_TP(3, false, true) : INSEM_TP(3, false, false);
  }
#undef INSEM_TP
  return hipGetLastError();
}

hipError_t launch_predictor(const DomDev& d, const StepCoef& sc, const StepClock& clk, hipStream_t s) {
  const int64_t n = d.ncomp * d.ng;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predictor, dim3(blocks_for(n)), dim3(kTB), 0, s, d, sc, clk);
  return hipGetLastError();
}

hipError_t launch_predictor_flat(const DomDev* d0, const DomDev* d1, const StepCoef& sc, hipStream_t s) {
  PredSpan sp[2];
  int64_t q[2] = {0, 0};
  const DomDev* dd[2] = {d0, d1};
  for (int k = 0; k < 2; ++k) {
    if (!dd[k]) continue;
    const DomDev& d = *dd[k];
    if ((reinterpret_cast<uintptr_t>(d.u) | reinterpret_cast<uintptr_t>(d.v) | reinterpret_cast<uintptr_t>(d.a)) & 15)
      return hipErrorInvalidValue;  // fields are 256-byte aligned allocations
    sp[k] = PredSpan{d.u, d.v, d.a, d.ncomp * d.ng};
    q[k] = (sp[k].n + 3) / 4;
  }
  const int64_t nb = (q[0] + q[1] + kTB - 1) / kTB;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predictor_flat, dim3(static_cast<unsigned>(nb)), dim3(kTB), 0, s, sp[0], sp[1], q[0], q[1],
                     sc.rdt, sc.hdt, sc.hdt2);
  return hipGetLastError();
}

hipError_t launch_assemble(const DomDev& d, hipStream_t s) {
  const int nb = blocks_for(d.ng);
  if (nb == 0) return hipSuccess;
  const bool w = d.t_w != nullptr;
#define INSEM_ASM(C)                                                           \
  if (w)                                                                       \
    hipLaunchKernelGGL((k_assemble<C, true>), dim3(nb), dim3(kTB), 0, s, d);  \
  else                                                                         \
    hipLaunchKernelGGL((k_assemble<C, false>), dim3(nb), dim3(kTB), 0, s, d);
  switch (d.ncomp) {
    case 1: INSEM_ASM(1) break;
    case 2: INSEM_ASM(2) break;
    case 3: INSEM_ASM(3) break;
    default: return hipErrorInvalidValue;
  }
#undef INSEM_ASM
  return hipGetLastError();
}

hipError_t launch_reduce(const double* partial, int n, int width, double* out, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kTB), 0, s, partial, n, width, out, accumulate);
  return hipGetLastError();
}

hipError_t launch_receivers(const ReceiverDev& r, const StepClock& clk, int pv_mode, double half_dt, hipStream_t s) {
  if (r.n_tasks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_receivers, dim3(r.n_tasks), dim3(kRecWG), 0, s, r, clk, pv_mode, half_dt);
  return hipGetLastError();
}

hipError_t launch_energy_row(const double* partial, int n, const StepClock& clk, double* rows, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_row, dim3(1), dim3(kTB), 0, s, partial, n, clk, rows);
  return hipGetLastError();
}

hipError_t launch_clock_add(int32_t* base, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_clock_add, dim3(1), dim3(64), 0, s, base, n);
  return hipGetLastError();
}

hipError_t launch_zero_points(float* x, int64_t ng, const int32_t* pts, int n, int comp_mask, int ncomp,
                              hipStream_t s) {
  if (n == 0 || comp_mask == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_points, dim3((n + 255) / 256), dim3(256), 0, s, x, ng, pts, n, comp_mask, ncomp);
  return hipGetLastError();
}

hipError_t launch_power_stats(const DomDev& d, float* x, double* partial, const double* state, int* n_blocks,
                              hipStream_t s) {
  const int nb = blocks_for(d.ng);
  *n_blocks = nb;
  hipLaunchKernelGGL(k_power_stats, dim3(nb), dim3(kTB), 0, s, d, x, partial, state);
  return hipGetLastError();
}

hipError_t launch_power_update(const double* out3, double* state, hipStream_t s) {
  hipLaunchKernelGGL(k_power_update, dim3(1), dim3(64), 0, s, out3, state);
  return hipGetLastError();
}

hipError_t launch_scale(float* x, int64_t n, const double* state, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(kTB), 0, s, x, n, state);
  return hipGetLastError();
}

hipError_t launch_mass_norm(const DomDev& d, const float* x, double* partial, int* n_blocks, hipStream_t s) {
  const int nb = blocks_for(d.ng);
  *n_blocks = nb;
  hipLaunchKernelGGL(k_mass_norm, dim3(nb), dim3(kTB), 0, s, d, x, partial);
  return hipGetLastError();
}

hipError_t launch_accel(const DomDev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_accel, dim3(blocks_for(d.ng)), dim3(kTB), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_coupling_power(const DomDev& d, const int32_t* pts, const int32_t* other, const float* cx,
                                 const float* cz, int n, int role, int pm, const float* src, int64_t src_ng,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_coupling_power, dim3((n + 255) / 256), dim3(256), 0, s, d, pts, other, cx, cz, n, role, pm,
                     src, src_ng);
  return hipGetLastError();
}

hipError_t launch_coupled_update(const double* nx, const double* ny, int it, double* state, hipStream_t s) {
  hipLaunchKernelGGL(k_coupled_update, dim3(1), dim3(64), 0, s, nx, ny, it, state);
  return hipGetLastError();
}

hipError_t launch_scale_from(float* x, const float* a, int64_t n, const double* state, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_from, dim3(blocks_for(n)), dim3(kTB), 0, s, x, a, n, state);
  return hipGetLastError();
}

hipError_t launch_dot(const float* a, const float* b, int64_t n, double* partial, int* n_blocks, hipStream_t s) {
  const int nb = blocks_for(n);
  *n_blocks = nb;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kTB), 0, s, a, b, n, partial);
  return hipGetLastError();
}

}  // namespace insem_hip
