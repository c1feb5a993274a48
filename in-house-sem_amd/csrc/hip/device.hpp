// Device-side data descriptors shared by the gfx950 kernels and the host
// code that launches them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

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

// Element i of an array that the kernels index with 32-bit offsets: a
// uniform base plus a 32-bit byte offset, which the load and store
// instructions take directly (no 64-bit address arithmetic per access).
// Every such array holds fewer than kMaxIndex32 elements; the host checks
// this when it builds a context.
constexpr int64_t kMaxIndex32 = int64_t{1} << 30;

template <class T>
__device__ __forceinline__ T* at32(T* p, uint32_t i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<B*>(p) + i * static_cast<uint32_t>(sizeof(T)));
}

// Value of a per-point quantity at slot k of element b, from the byte
// offsets b4 = 4 b and slot4 = 4 (b * npe + k): one select per quantity.
__device__ __forceinline__ float ld(const PF& f, uint32_t b4, uint32_t slot4) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(f.p) + (f.sk ? slot4 : b4));
}

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

// Grouped points of a domain: regular points whose contributions all come
// from element slots of their own (no constraint weights), numbered in
// groups of equal contribution count: group q holds the points
// [start[q], start[q] + size[q]), each with count[q] contributions.  The
// contributions lie in rows of stride = end - begin values: component c,
// rank r (the slot's place in the point's ascending slot order) of grouped
// point g at contrib[(c * rows + r) * stride + (g - begin)], rows = the
// largest count.  A slot so finds its place from its rank alone, and the
// tail reads every row in sequence.
constexpr int kMaxGroups = 7;  // contribution counts 2 .. 8
struct Groups {
  int n = 0;
  int count[kMaxGroups] = {};
  int64_t start[kMaxGroups] = {};
  int64_t size[kMaxGroups] = {};
  int rows = 0;
  int64_t begin = 0, end = 0;  // [begin, end): all grouped points
};

// Contribution count of grouped point g: unrolled selects over the
// (uniform) table, no indexed loads.
__device__ __forceinline__ int group_count(const Groups& G, uint32_t g) {
  int k = G.count[0];
#pragma unroll
  for (int q = 1; q < kMaxGroups; ++q) {
    const bool in = q < G.n && g >= static_cast<uint32_t>(G.start[q]);
    k = in ? G.count[q] : k;
  }
  return k;
}

// Slot codes (16 bits per element slot; 0 for any other slot).
// Grouped point: the flag and the slot's rank among the point's
// contributions.  Batch-local point (all of its slots in one batch
// of one force launch): the flag, whether the slot is the point's first in
// slot order (its leader), the point's slot count - 1 and the slot's place z
// in the batch's force buffer (the point's slots take consecutive places in
// slot order).
constexpr uint16_t kSlotGrouped = 0x8000;
constexpr uint16_t kSlotLocal = 0x4000;
constexpr uint16_t kSlotLeader = 0x1000;
__host__ __device__ constexpr uint16_t slot_code(int rank) { return static_cast<uint16_t>(kSlotGrouped | rank); }
// z < 512 (9 bits), count 2 .. 8 (3 bits).
__host__ __device__ constexpr uint16_t local_code(bool leader, int count, int z) {
  return static_cast<uint16_t>(kSlotLocal | (leader ? kSlotLeader : 0) | ((count - 1) << 9) | z);
}
__host__ __device__ constexpr int local_place(int code) { return code & 0x1FF; }
__host__ __device__ constexpr int local_count(int code) { return ((code >> 9) & 7) + 1; }

// Time-step coefficients of one model of a batch, computed as a single
// model's are (float(dt), float(dt/2), float(dt^2/2), dt/2, dt^2/8).
struct ModelCoef {
  float rdt = 0.0f, hdt = 0.0f, hdt2 = 0.0f, pad = 0.0f;
  double h_damp = 0.0;   // dt/2: Stacey implicit weight
  double cc = 0.0;       // dt^2/8: kinetic energy
  double half_dt = 0.0;  // particle-velocity trapezoid
  int32_t n_steps = 0;   // the model's last step
  int32_t pad2 = 0;
};

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
  // Time step with the update fused in (StepParams::fuse): the points
  // [0, n_int) have unit-weight unconstrained slots of their own in one
  // batch only, and no boundary or source work.  A point of one slot takes
  // that slot's force as its assembled force; a batch-local point's slots
  // leave their forces in LDS and its first slot sums them.  That work-item
  // finishes the point's update (M^-1, corrector, energy partials, next
  // predictor) instead of writing f_local.
  int64_t n_int = 0;
  float* uw = nullptr;
  float* v = nullptr;
  float* a = nullptr;
  const float* im = nullptr;
  const float* mass = nullptr;
  // Slots of grouped points write their force into the group rows
  // instead of f_local (fused steps only).
  const uint16_t* code = nullptr;  // [nb][npe] slot codes (slot_code, local_code)
  int has_local = 0;               // some slot of the block is batch-local
  // Group rows: first grouped point, row stride, rows per component.
  int gbegin = 0, gstride = 0, grows = 0;
  float* contrib = nullptr;
  // Batches: model of each block element and the models' coefficients.
  const int32_t* emodel = nullptr;
  const ModelCoef* models = nullptr;
  const int32_t* batch_last = nullptr;  // per batch: the last step of its models (a batch's launch skips it after)
  const int32_t* batch_model = nullptr;  // per batch: its one model, or -1 (elements of several models)
};

// Per-launch parameters of the element-force kernels.
struct StepParams {
  int fuse = 0;  // a time step: update owned points, write grouped forces
  int corr = 0, pred = 0;
  float rdt = 0.0f, hdt = 0.0f, hdt2 = 0.0f;  // float(dt), float(dt/2), float(dt^2/2)
  double cc = 0.0;                            // dt^2 / 8 (kinetic energy)
  StepClock clk;
  double* partial = nullptr;  // [work-group][2]: u.f, kinetic (energy steps)
};

// Global-point data of one domain on the device, in the backend's point
// numbering: points that need boundary, coupling, source or receiver work
// ("special") are numbered after all others, so the bulk update runs
// branch-free over [0, n_reg).
struct DomDev {
  int ncomp = 1;
  int npe = 1;
  int64_t ng = 0;
  int64_t n_int = 0;  // [0, n_int): updated by the element kernels in a time step
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
  Groups grp;                       // grouped points (fused steps)
  const float* contrib = nullptr;
  // Batches: model of each point and the models' coefficients (null for a
  // single model, whose coefficients come with the launch).
  const int32_t* pmodel = nullptr;
  const ModelCoef* models = nullptr;
  const int32_t* chunk_model = nullptr;  // per 256 points: their one model, or -1 (points of several models)
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

}  // namespace insem_hip
