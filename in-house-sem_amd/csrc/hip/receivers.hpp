// Receiver sampling on the device (reference: solver.cpp:780-811, 945-958),
// shared by the standalone receiver launch and the element-force launch of
// the next step, which carries the receivers of the previous one.
//
// One work-group per task (receiver quantity): the sums run over the
// receiver's points in their order, accumulated in double from the float
// fields as the reference does, and the particle-velocity trapezoid too.
#pragma once

#include "../../../include/insem_hip.h"
#include "device.hpp"
#include "launch.hpp"

namespace insem_hip {

// Sums w[q] * x[stride + pts[q]] over q in [p0, p1) in ascending q, in
// double: the NT work-items stage the operands in LDS, work-item 0 adds them
// in order (the reference's serial order, without a chain of dependent
// global loads).  Every work-item of the group must call it.
template <int NT>
__device__ double ordered_dot(const float* w, const float* x, const int32_t* pts, int64_t stride, int32_t p0,
                              int32_t p1, double* sw, float* sx) {
  double s = 0.0;
  for (int32_t base = p0; base < p1; base += NT) {
    const int32_t q = base + static_cast<int32_t>(threadIdx.x);
    if (q < p1) {
      sw[threadIdx.x] = w[q];
      sx[threadIdx.x] = x[stride + pts[q]];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int32_t n = p1 - base < NT ? p1 - base : NT;
      for (int32_t i = 0; i < n; ++i) s += sw[i] * static_cast<double>(sx[i]);
    }
    __syncthreads();
  }
  return s;
}

// Task t at `step`: records the sample (on recording steps) and advances
// the particle velocity (pv_mode 0: take prev from the current field; 1:
// integrate).  sw, sx: NT doubles and NT floats of LDS.
template <int NT>
__device__ void receiver_task(const ReceiverDev& r, int t, const StepClock& clk, int step, int pv_mode,
                              double half_dt, double* sw, float* sx) {
  if (r.models) {
    // a batch: the task's own model's step size, nothing past its last step
    const ModelCoef& m = r.models[r.tmodel[t]];
    if (step > m.n_steps) return;
    half_dt = m.half_dt;
  }
  const int sample =
      step % clk.record_every == 0 && step / clk.record_every < clk.max_samples ? step / clk.record_every : -1;
  const int dom = r.domain[t];
  const int kind = r.kind[t];
  const int64_t nsp = r.n_sp[dom];
  const int32_t p0 = r.pt_off[t], p1 = r.pt_off[t + 1];
  double* row = sample >= 0 ? r.traces + static_cast<int64_t>(sample) * r.n_cols + r.col[t] : nullptr;
  if (kind == INSEM_Q_PARTICLE_VELOCITY) {
    const float* u = r.su[dom];
    const double gx = ordered_dot<NT>(r.d_dx, u, r.pts, 0, p0, p1, sw, sx);
    const double gz = ordered_dot<NT>(r.d_dz, u, r.pts, 0, p0, p1, sw, sx);
    if (threadIdx.x == 0) {
      const double ax = -gx / r.rho[t], az = -gz / r.rho[t];
      double* pv = r.pv + 4 * t;  // prev_ax, prev_az, vel_x, vel_z
      if (pv_mode == 1) {
        pv[2] += half_dt * (pv[0] + ax);
        pv[3] += half_dt * (pv[1] + az);
      }
      pv[0] = ax;
      pv[1] = az;
      if (row) {
        row[0] = pv[2];
        row[1] = pv[3];
      }
    }
  } else if (row) {
    const float* field =
        kind == INSEM_Q_VELOCITY ? r.sv[dom] : kind == INSEM_Q_ACCELERATION ? r.sa[dom] : r.su[dom];
    for (int c = 0; c < r.ncomp[t]; ++c) {
      const double s = ordered_dot<NT>(r.value, field, r.pts, c * nsp, p0, p1, sw, sx);
      if (threadIdx.x == 0) row[c] = s;
    }
  }
}

}  // namespace insem_hip
