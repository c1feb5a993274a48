// Device-side data descriptors shared by the gfx950 kernels and the host
// code that launches them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace insem_hip {

// Physics families of the element kernels (reference: the six kernel
// templates of cpp/src/cpu/kernels.cpp:64-431).
enum Phys : int {
  kAcousticCart = 0,
  kElasticCart = 1,
  kAcousticAxi0 = 2,
  kElasticAxi0 = 3,
  kAcousticAxi1 = 4,
  kElasticAxi1 = 5,
};

__host__ __device__ constexpr int phys_ncomp(int phys) {
  return phys == kElasticCart || phys == kElasticAxi0 ? 2 : phys == kElasticAxi1 ? 3 : 1;
}

// A per-point quantity of a block, stored either per point ([b][k]: sb = npe,
// sk = 1) or, when it is bitwise constant over every element of the block,
// once per element ([b]: sb = 1, sk = 0).  Coefficients are constant per
// element in every model (one material per leaf), and so are some metric
// components (e.g. xi_z = 0 on elements with vertical sides); storing them
// once per element removes their HBM traffic without changing any value.
struct PF {
  const float* p = nullptr;
  int sb = 0;
  int sk = 0;
};

__device__ __forceinline__ float ld(const PF& f, int b, int k) {
  return f.p[static_cast<size_t>(b) * f.sb + static_cast<size_t>(k) * f.sk];
}

// Arguments of one element-force launch (one block of one domain).
struct ForceArgs {
  int nb = 0;
  // Production path: gather from the global field through l2g ([nb][npe],
  // new point numbering, < 0: constraint row -1 - l2g).
  const int32_t* l2g = nullptr;
  const int32_t* row_off = nullptr;
  const int32_t* row_dofs = nullptr;
  const float* row_w = nullptr;
  const float* u = nullptr;
  int64_t ng = 0;
  // Operator seam: element-local input in the reference layout.
  const float* ul = nullptr;
  const int32_t* elem = nullptr;  // domain element of each block element
  // Output: f_local at element (out_pos0 + b), or at elem[b] when out_ref.
  float* fl = nullptr;
  int out_pos0 = 0;
  int out_ref = 0;
  PF m[5];  // xi_x, xi_z, eta_x, eta_z, jac (xi_r ... on axisymmetric blocks)
  PF c0, c1, wr, hoop, r;
  const uint8_t* axial = nullptr;
  const float* D = nullptr;
  const float* Dj = nullptr;
  const float* w2 = nullptr;
};

// Global-point data of one domain on the device, in the backend's point
// numbering: points that need boundary, coupling, source or receiver work
// ("special") are numbered after all others, so the bulk update runs
// branch-free over [0, n_reg).
struct DomDev {
  int ncomp = 1;
  int npe = 1;
  int64_t ng = 0;
  int64_t n_reg = 0;
  float* u = nullptr;
  float* v = nullptr;
  float* a = nullptr;
  float* f = nullptr;  // assembled forces (power iteration and operator seam)
  const float* im = nullptr;
  const float* mass = nullptr;
  const int32_t* t_off = nullptr;
  const int32_t* t_addr = nullptr;  // f_local address of each contribution
  const float* t_w = nullptr;       // null without constraints
  const float* fl = nullptr;
};

// Special-point data of one domain ([0, n_sp) = points n_reg .. ng - 1).
struct SpecialDev {
  int64_t n_sp = 0;
  const int32_t* abs_idx = nullptr;    // Stacey matrix index or -1
  const float* abs_coef = nullptr;     // [n_abs][ncomp * ncomp]
  const uint8_t* flags = nullptr;      // bit 0 Dirichlet, bit 1 axis
  int axis_mask = 0;                   // bit c: component c vanishes on the axis
  const int32_t* src_off = nullptr;    // [n_sp + 1]
  const int32_t* src_id = nullptr;     // source index of each entry
  const float* src_coef = nullptr;     // [entry][ncomp]
  const float* amps = nullptr;         // [n_sources][n_steps + 1]
  int amp_stride = 0;
  const int32_t* cpl_idx = nullptr;    // coupling entry or -1
  const int32_t* cpl_other = nullptr;  // point of the other domain (its numbering)
  const float* cpl_cx = nullptr;
  const float* cpl_cz = nullptr;
  int cpl_role = 0;                    // 1 solid (reads fluid u), 2 fluid (reads solid a)
  int pm = 0;                          // azimuthal order 1
  const float* other = nullptr;        // fluid u or solid a
  int64_t other_ng = 0;
  float* su = nullptr;                 // [ncomp][n_sp]: pre-predictor values
  float* sv = nullptr;
  float* sa = nullptr;
};

// Step number of a launch: *base (the step before the chunk, on the device)
// + offset (baked into the launch).  Replayed graphs advance base only.
struct StepClock {
  const int32_t* base = nullptr;
  int offset = 0;
  int energy_every = 0;
  int record_every = 1;
  int max_samples = 0;
};

__device__ __forceinline__ int clock_step(const StepClock& k) { return *k.base + k.offset; }
__device__ __forceinline__ bool clock_energy(const StepClock& k, int step) {
  return k.energy_every > 0 && step % k.energy_every == 0;
}

}  // namespace insem_hip
