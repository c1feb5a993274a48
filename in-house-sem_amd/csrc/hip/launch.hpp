// Host-side launchers of the gfx950 kernels.
#pragma once

#include "device.hpp"

namespace insem_hip {

// Receiver tasks (one per receiver quantity) on the device: interpolation
// weights over the receiver's points, which index the side buffers of the
// special points (the pre-predictor u, v, a of each step).
struct ReceiverDev {
  int n_tasks = 0;
  const int32_t* domain = nullptr;  // per task (receiver quantity): side-buffer slot 0/1
  const int32_t* kind = nullptr;    // INSEM_Q_*
  const int32_t* ncomp = nullptr;   // raw device components
  const int32_t* col = nullptr;     // first column in a trace row
  const int32_t* pt_off = nullptr;  // [n_tasks + 1]
  const int32_t* pts = nullptr;     // side-buffer index of each point
  const float* value = nullptr;
  const float* d_dx = nullptr;
  const float* d_dz = nullptr;
  const double* rho = nullptr;
  double* pv = nullptr;  // [task][4]: prev_ax, prev_az, vel_x, vel_z
  int n_cols = 0;
  double* traces = nullptr;  // [sample][n_cols]
  const float* su[2] = {nullptr, nullptr};
  const float* sv[2] = {nullptr, nullptr};
  const float* sa[2] = {nullptr, nullptr};
  int64_t n_sp[2] = {0, 0};
  // Batches: model of each task and the models' coefficients.
  const int32_t* tmodel = nullptr;
  const ModelCoef* models = nullptr;
};

// Element forces of one block (gather = read the global field through l2g;
// otherwise read the element-local field in the reference layout).  With
// p.fuse (gather only), the points [0, a.n_int) are updated in the
// kernel, grouped points receive their forces in the group rows and the
// energy partials go to p.partial[work-group][2]; *grid (optional) receives
// the launch's work-group count.
hipError_t launch_forces(int ngll, int phys, bool rect, bool gather, const ForceArgs& a, const StepParams& p,
                         hipStream_t s, int* grid = nullptr, const ReceiverDev* rec = nullptr, int n_rec = 0,
                         double half_dt = 0.0);
// With rec (gather only), the launch's first n_rec work-groups sample the
// receivers at the step before the clock's (integrating the particle
// velocity), ahead of the element work-groups.
// Work-groups of one force launch over nb elements.
int force_blocks(int ngll, int nb);
// Elements per batch (one work-group's share) of a force launch.
int force_batch_elements(int ngll);
// A time step's force launch over an elastic block a and an acoustic block
// b together (fused steps only), for the physics pairs of a coupled model:
// (elastic, acoustic) Cartesian, axisymmetric m = 0 or m = 1.  Receivers as
// in launch_forces.  The launch has at most 14 work-groups of padding more
// than the two blocks' force_blocks and n_rec.
bool forces_pair_supported(int pa, bool ra, int pb, bool rb);
// variant: -1 the tuned one (tuned_force_variant), else level + kForceLevels
// * sched of the autotuner's space (level: registers budgeted for the
// default waves per SIMD, at most 6, at most 4; sched: LLVM's max-ilp or
// default scheduling strategy).  Variants other than 0 exist for orders 4 to
// 6 (ngll 5 to 7); elsewhere every variant runs variant 0.  All variants give
// the same bits.
constexpr int kForceLevels = 3;
constexpr int kForceVariants = 2 * kForceLevels;
hipError_t launch_forces_pair(int ngll, int pa, bool ra, int pb, bool rb, const ForceArgs& a, const ForceArgs& b,
                              const StepParams& p, hipStream_t s, int* grid, const ReceiverDev* rec, int n_rec,
                              double half_dt, int variant = -1);
int tuned_force_variant(int ngll, int pa, bool ra, int pb, bool rb);
// The tuned steps per captured chunk of a model with this paired launch (0:
// none recorded).
int tuned_chunk_steps(int ngll, int pa, bool ra, int pb, bool rb);
// The variants of one scheduling strategy (level as above).
hipError_t launch_forces_pair_max_ilp(int ngll, int level, int pa, bool ra, int pb, bool rb, const ForceArgs& a,
                                      const ForceArgs& b, const StepParams& p, hipStream_t s, int* grid,
                                      const ReceiverDev* rec, int n_rec, double half_dt);
hipError_t launch_forces_pair_default_sched(int ngll, int level, int pa, bool ra, int pb, bool rb,
                                            const ForceArgs& a, const ForceArgs& b, const StepParams& p,
                                            hipStream_t s, int* grid, const ReceiverDev* rec, int n_rec,
                                            double half_dt);

struct StepCoef {
  float rdt = 0.0f, hdt = 0.0f, hdt2 = 0.0f;  // float(dt), float(dt/2), float(dt^2/2)
  double h_damp = 0.0;                        // dt/2, or 0 for the explicit damping at t = 0
  double cc = 0.0;                            // dt^2 / 8 (kinetic energy)
};

int tail_blocks(int64_t n);
// Tail of a step on one domain, one launch: a = M^-1 f from the assembled
// forces (grouped points from the group rows when `grouped`, the other
// regular points through the transpose CSR), sources, coupling, Stacey,
// Dirichlet and axis conditions at the special points (whose pre-predictor
// values go to the side buffers), optional corrector, energy partials
// ([blocks][2]: u.f, kinetic) and the next step's predictor.  The step
// number and whether it is an energy step come from the clock.
hipError_t launch_tail(const DomDev& d, const SpecialDev& sp, const StepCoef& sc, int corrector, int predictor,
                       const StepClock& clk, bool grouped, double* partial, int* n_blocks, hipStream_t s);
// Fused steps of a coupled model: the tails of both domains in one launch
// (each interface point's fluid partner is updated right after it, by the
// same work-item, with its new acceleration).
hipError_t launch_tail_pair(const DomDev& solid, const SpecialDev& solid_sp, const DomDev& fluid,
                            const SpecialDev& fluid_sp, const StepCoef& sc, int corrector, int predictor,
                            const StepClock& clk, double* partial, int* n_blocks, hipStream_t s);
// rows[step / energy_every] = sums of [n][2] partials, on energy steps only.
hipError_t launch_energy_row(const double* partial, int n, const StepClock& clk, double* rows, hipStream_t s);
// *base += n (end of a chunk of n steps).
hipError_t launch_clock_add(int32_t* base, int n, hipStream_t s);
// u += rdt v + hdt2 a; v += hdt a over all points, for the step after the
// clock's (per-model coefficients on a batch).
hipError_t launch_predictor(const DomDev& d, const StepCoef& sc, const StepClock& clk, hipStream_t s);
// The predictor of one model (not a batch) over both domains in one launch
// (d1 may be null).
hipError_t launch_predictor_flat(const DomDev* d0, const DomDev* d1, const StepCoef& sc, hipStream_t s);
// f = assembly of f_local over all points (no update).
hipError_t launch_assemble(const DomDev& d, hipStream_t s);
// Fixed-order sum of n rows of `width` doubles into out (+= when accumulate).
hipError_t launch_reduce(const double* partial, int n, int width, double* out, int accumulate, hipStream_t s);

// Records the sample of the clock's step (when it is a recording step) and
// advances the particle velocity (pv_mode 0: take prev from the current
// field; 1: integrate).
hipError_t launch_receivers(const ReceiverDev& r, const StepClock& clk, int pv_mode, double half_dt, hipStream_t s);

// Power iteration pieces (FP64 reductions).  state = {lambda, done, scale}.
hipError_t launch_zero_points(float* x, int64_t ng, const int32_t* pts, int n, int comp_mask, int ncomp,
                              hipStream_t s);
hipError_t launch_power_stats(const DomDev& d, float* x, double* partial, const double* state, int* n_blocks,
                              hipStream_t s);
hipError_t launch_power_update(const double* out3, double* state, hipStream_t s);
hipError_t launch_scale(float* x, int64_t n, const double* state, hipStream_t s);
hipError_t launch_mass_norm(const DomDev& d, const float* x, double* partial, int* n_blocks, hipStream_t s);
hipError_t launch_accel(const DomDev& d, hipStream_t s);
hipError_t launch_coupling_power(const DomDev& d, const int32_t* pts, const int32_t* other, const float* cx,
                                 const float* cz, int n, int role, int pm, const float* src, int64_t src_ng,
                                 hipStream_t s);
hipError_t launch_coupled_update(const double* nx, const double* ny, int it, double* state, hipStream_t s);
hipError_t launch_scale_from(float* x, const float* a, int64_t n, const double* state, hipStream_t s);
hipError_t launch_dot(const float* a, const float* b, int64_t n, double* partial, int* n_blocks, hipStream_t s);

}  // namespace insem_hip
