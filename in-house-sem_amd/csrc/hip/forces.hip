// Element internal-force kernels for gfx950: f_local = -K u_local of one
// element block, with the gather from the global field fused in.
//
// Mapping: one work-item per element point, 256 work-items per work-group
// holding floor(256 / npe) whole elements (10 at order 4), a "batch".  The
// gathered element fields and the reference fluxes are staged in LDS; each
// work-item keeps the rows and columns of the derivative matrices it needs
// in registers.  Every per-point arithmetic expression and every summation
// follows the reference kernels (cpp/src/cpu/kernels.cpp:19-431) in the same
// order, and the library is compiled with -ffp-contract=off, so the result
// is bitwise the reference's float result.
//
// Global loads are issued in the order their results are needed (the shared
// matrices, the slot's point, the per-point operands, then everything that
// depends on the point) and without branches, so that each wait covers only
// what it must.  Work-group ids are remapped so that each of the 8 XCDs
// processes one contiguous range of elements (neighbouring elements share
// global points, which then stay in that XCD's L2).
//
// In a time step (StepParams::fuse) the slot's force is consumed directly:
// a point owned by this slot alone is updated in place, a grouped point's
// slot writes its force into the group rows, and only the remaining
// slots (special points, constrained slots, other points) write f_local.
// The step's first force launch carries the previous step's receiver tasks
// in its first work-groups.

#include "device.hpp"
#include "launch.hpp"
#include "receivers.hpp"

namespace insem_hip {
namespace {

constexpr int kWG = 256;

__device__ __forceinline__ int xcd_contiguous(int bid, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int x = bid & 7, s = bid >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}

template <int NG>
__device__ __forceinline__ void deriv(const float* xr, const float* yr, const float* U, int i, int j, float& u_xi,
                                      float& u_eta) {
  float a = 0.0f, b = 0.0f;
#pragma unroll
  for (int l = 0; l < NG; ++l) {
    a += xr[l] * U[j * NG + l];
    b += yr[l] * U[l * NG + i];
  }
  u_xi = a;
  u_eta = b;
}

template <int NG>
__device__ __forceinline__ float wdiv(const float* xc, const float* yc, const float* FX, const float* FE, int i, int j) {
  float s = 0.0f;
#pragma unroll
  for (int l = 0; l < NG; ++l) {
    s += xc[l] * FX[j * NG + l];
    s += yc[l] * FE[l * NG + i];
  }
  return s;
}

__device__ __forceinline__ void pgrad(float a_xi, float a_eta, float xi_x, float xi_z, float eta_x, float eta_z,
                                      float& a_x, float& a_z) {
  a_x = a_xi * xi_x + a_eta * eta_x;
  a_z = a_xi * xi_z + a_eta * eta_z;
}

__device__ __forceinline__ void rflux(float wj, float qx, float qz, float xi_x, float xi_z, float eta_x, float eta_z,
                                      float& f_xi, float& f_eta) {
  f_xi = wj * (qx * xi_x + qz * xi_z);
  f_eta = wj * (qx * eta_x + qz * eta_z);
}

// Axis spread of an axial element at point (i, j) (kernels.cpp:195-204,
// the updates that reach this point, in the reference's order).
template <int NG>
__device__ __forceinline__ float spread(float f, int i, int j, const float* g, const float* xr0, const float* er0,
                                        const float* DxRow0, const float* Dy) {
  if (i > 0) return f - g[j] * xr0[j] * DxRow0[i];
#pragma unroll
  for (int jq = 0; jq < NG; ++jq) {
    if (jq == j) f -= g[j] * xr0[j] * DxRow0[0];
    f -= g[jq] * er0[jq] * Dy[jq * NG + j];
  }
  return f;
}

template <int PHYS>
struct Shape {
  static constexpr int C = phys_ncomp(PHYS);
  static constexpr bool AXI = PHYS >= kAcousticAxi0;
  static constexpr int NAX = PHYS == kElasticAxi1 ? 2 : 1;
};

// LDS tiles of one work-group.
template <int NG, int PHYS>
struct Tiles {
  static constexpr int NPE = NG * NG, EPB = kWG / NPE, C = Shape<PHYS>::C;
  static constexpr bool AXI = Shape<PHYS>::AXI;
  static constexpr int NAX = Shape<PHYS>::NAX;
  float D[NPE], Dj[NPE], W2[NPE];
  float u[EPB * C * NPE];
  float f[EPB * 2 * C * NPE];
  float fo[kWG * C];  // forces of batch-local slots, [z][c]
  float g[AXI ? EPB * NAX * NG : 1];
  float xr[AXI ? EPB * NG : 1];
  float er[AXI ? EPB * NG : 1];
};

// Per-point operands of an element slot: metric (xi_x, xi_z, eta_x, eta_z,
// jac; xi_r ... on axisymmetric blocks), coefficients and the axisymmetric
// factors.
// (Set by load_ops for active slots only: an inactive lane holds no values.)
struct Ops {
  float q0, q1, q2, q3, q4, k0, k1;
  float qw, qh, qr;
  bool axial = false;
};

template <int PHYS, bool RECT>
__device__ __forceinline__ void load_ops(const ForceArgs& A, uint32_t b, uint32_t slot, Ops& o) {
  constexpr int C = Shape<PHYS>::C;
  constexpr bool AXI = Shape<PHYS>::AXI;
  const uint32_t b4 = b * 4u, s4 = slot * 4u;
  o.q0 = ld(A.m[0], b4, s4);
  o.q3 = ld(A.m[3], b4, s4);
  o.q4 = AXI ? 0.0f : ld(A.m[4], b4, s4);
  o.q1 = RECT ? 0.0f : ld(A.m[1], b4, s4);
  o.q2 = RECT ? 0.0f : ld(A.m[2], b4, s4);
  o.k0 = ld(A.c0, b4, s4);
  o.k1 = C >= 2 ? ld(A.c1, b4, s4) : 0.0f;
  o.qw = AXI ? ld(A.wr, b4, s4) : 0.0f;
  o.qr = AXI ? ld(A.r, b4, s4) : 0.0f;
  o.qh = AXI && PHYS != kAcousticAxi0 ? ld(A.hoop, b4, s4) : 0.0f;
  if constexpr (AXI) o.axial = *at32(A.axial, b) != 0;
}

// Derivative-matrix rows (for the gradient) and columns (for the weak
// divergence) of point (i, j), GLJ in xi on axial elements.  Axisymmetric
// kernels read the columns after the flux barrier, so that they are not
// live in registers through the flux computation (which raises their
// occupancy); the Cartesian ones, at full occupancy either way, read them
// with the rows.
template <int NG, int PHYS>
__device__ __forceinline__ void d_rows(const Tiles<NG, PHYS>& T, bool axial, int i, int j, float* xr, float* yr) {
  const float* Dx = axial ? T.Dj : T.D;
#pragma unroll
  for (int l = 0; l < NG; ++l) {
    xr[l] = Dx[i * NG + l];
    yr[l] = T.D[j * NG + l];
  }
}

template <int NG, int PHYS>
__device__ __forceinline__ void d_cols(const Tiles<NG, PHYS>& T, bool axial, int i, int j, float* xc, float* yc) {
  const float* Dx = axial ? T.Dj : T.D;
#pragma unroll
  for (int l = 0; l < NG; ++l) {
    xc[l] = Dx[l * NG + i];
    yc[l] = T.D[l * NG + j];
  }
}

// Reference fluxes of point k of element le (from the field tile) into the
// flux tile, and the direct (undifferentiated) terms of the point.
template <int NG, int PHYS, bool RECT>
__device__ __forceinline__ void element_flux(Tiles<NG, PHYS>& T, const Ops& o, int le, int k, int i, int j,
                                             const float* xr, const float* yr, float& direct0, float& direct1) {
  constexpr int NPE = NG * NG;
  constexpr int C = Shape<PHYS>::C;
  constexpr int NAX = Shape<PHYS>::NAX;
  const float* U = T.u + le * C * NPE;
  float* F = T.f + le * 2 * C * NPE;
  const bool axial = o.axial;
  direct0 = 0.0f;
  direct1 = 0.0f;
  if constexpr (PHYS == kElasticCart) {
    float ux_xi, ux_eta, uz_xi, uz_eta;
    deriv<NG>(xr, yr, U, i, j, ux_xi, ux_eta);
    deriv<NG>(xr, yr, U + NPE, i, j, uz_xi, uz_eta);
    float xi_x, xi_z = 0.0f, eta_x = 0.0f, eta_z, jac;
    xi_x = o.q0;
    eta_z = o.q3;
    jac = o.q4;
    if constexpr (!RECT) {
      xi_z = o.q1;
      eta_x = o.q2;
    }
    float ux_x, ux_z, uz_x, uz_z;
    if constexpr (RECT) {
      ux_x = ux_xi * xi_x;
      ux_z = ux_eta * eta_z;
      uz_x = uz_xi * xi_x;
      uz_z = uz_eta * eta_z;
    } else {
      pgrad(ux_xi, ux_eta, xi_x, xi_z, eta_x, eta_z, ux_x, ux_z);
      pgrad(uz_xi, uz_eta, xi_x, xi_z, eta_x, eta_z, uz_x, uz_z);
    }
    const float lambda = o.k0, mu = o.k1;
    const float div = ux_x + uz_z;
    const float sxx = lambda * div + 2.0f * mu * ux_x;
    const float szz = lambda * div + 2.0f * mu * uz_z;
    const float sxz = mu * (ux_z + uz_x);
    const float wj = T.W2[k] * jac;
    float f1x, f1z, f2x, f2z;
    if constexpr (RECT) {
      f1x = wj * sxx * xi_x;
      f1z = wj * sxz * xi_x;
      f2x = wj * sxz * eta_z;
      f2z = wj * szz * eta_z;
    } else {
      rflux(wj, sxx, sxz, xi_x, xi_z, eta_x, eta_z, f1x, f2x);
      rflux(wj, sxz, szz, xi_x, xi_z, eta_x, eta_z, f1z, f2z);
    }
    F[0 * NPE + k] = f1x;
    F[1 * NPE + k] = f2x;
    F[2 * NPE + k] = f1z;
    F[3 * NPE + k] = f2z;
  } else if constexpr (PHYS == kAcousticCart) {
    float p_xi, p_eta;
    deriv<NG>(xr, yr, U, i, j, p_xi, p_eta);
    float xi_x, xi_z = 0.0f, eta_x = 0.0f, eta_z, jac;
    xi_x = o.q0;
    eta_z = o.q3;
    jac = o.q4;
    if constexpr (!RECT) {
      xi_z = o.q1;
      eta_x = o.q2;
    }
    const float c = T.W2[k] * jac * o.k0;
    float f1, f2;
    if constexpr (RECT) {
      f1 = c * p_xi * xi_x * xi_x;
      f2 = c * p_eta * eta_z * eta_z;
    } else {
      float p_x, p_z;
      pgrad(p_xi, p_eta, xi_x, xi_z, eta_x, eta_z, p_x, p_z);
      rflux(c, p_x, p_z, xi_x, xi_z, eta_x, eta_z, f1, f2);
    }
    F[k] = f1;
    F[NPE + k] = f2;
  } else {
    const float xi_r = o.q0, xi_z = o.q1;
    const float eta_r = o.q2, eta_z = o.q3;
    if (axial && i == 0) {
      T.xr[le * NG + j] = xi_r;
      T.er[le * NG + j] = eta_r;
    }
    if constexpr (PHYS == kAcousticAxi0) {
      float p_xi, p_eta, p_r, p_z;
      deriv<NG>(xr, yr, U, i, j, p_xi, p_eta);
      pgrad(p_xi, p_eta, xi_r, xi_z, eta_r, eta_z, p_r, p_z);
      float f1, f2;
      rflux(o.qw * o.k0, p_r, p_z, xi_r, xi_z, eta_r, eta_z, f1, f2);
      F[k] = f1;
      F[NPE + k] = f2;
    } else if constexpr (PHYS == kAcousticAxi1) {
      float p_xi, p_eta, p_r, p_z;
      deriv<NG>(xr, yr, U, i, j, p_xi, p_eta);
      pgrad(p_xi, p_eta, xi_r, xi_z, eta_r, eta_z, p_r, p_z);
      const float inv_rho = o.k0;
      const float w = o.qw;
      float f1, f2;
      rflux(w * inv_rho, p_r, p_z, xi_r, xi_z, eta_r, eta_z, f1, f2);
      F[k] = f1;
      F[NPE + k] = f2;
      if (axial && i == 0) {
        direct0 = 0.0f;
        T.g[le * NG + j] = w * inv_rho * p_r;
      } else {
        direct0 = o.qh * inv_rho * U[k] / o.qr;
      }
    } else if constexpr (PHYS == kElasticAxi0) {
      float ur_xi, ur_eta, uz_xi, uz_eta;
      deriv<NG>(xr, yr, U, i, j, ur_xi, ur_eta);
      deriv<NG>(xr, yr, U + NPE, i, j, uz_xi, uz_eta);
      float ur_r, ur_z, uz_r, uz_z;
      pgrad(ur_xi, ur_eta, xi_r, xi_z, eta_r, eta_z, ur_r, ur_z);
      pgrad(uz_xi, uz_eta, xi_r, xi_z, eta_r, eta_z, uz_r, uz_z);
      const float r = o.qr;
      const bool on_axis = r == 0.0f;
      const float e_tt = on_axis ? ur_r : U[k] / r;
      const float lambda = o.k0, mu = o.k1;
      const float tr = ur_r + uz_z + e_tt;
      const float srr = lambda * tr + 2 * mu * ur_r;
      const float szz = lambda * tr + 2 * mu * uz_z;
      const float stt = lambda * tr + 2 * mu * e_tt;
      const float srz = mu * (ur_z + uz_r);
      const float w = o.qw;
      float f1r, f2r, f1z, f2z;
      rflux(w, srr, srz, xi_r, xi_z, eta_r, eta_z, f1r, f2r);
      rflux(w, srz, szz, xi_r, xi_z, eta_r, eta_z, f1z, f2z);
      F[0 * NPE + k] = f1r;
      F[1 * NPE + k] = f2r;
      F[2 * NPE + k] = f1z;
      F[3 * NPE + k] = f2z;
      if (axial && i == 0) {
        direct0 = 0.0f;
        T.g[le * NG + j] = w * stt;
      } else {
        direct0 = o.qh * stt;
      }
    } else {  // kElasticAxi1
      const float s2 = 1.41421356237309504880f;
      const float is2 = 0.70710678118654752440f;
      float up_xi, up_eta, um_xi, um_eta, uz_xi, uz_eta;
      deriv<NG>(xr, yr, U, i, j, up_xi, up_eta);
      deriv<NG>(xr, yr, U + NPE, i, j, um_xi, um_eta);
      deriv<NG>(xr, yr, U + 2 * NPE, i, j, uz_xi, uz_eta);
      float up_r, up_z, um_r, um_z, uz_r, uz_z;
      pgrad(up_xi, up_eta, xi_r, xi_z, eta_r, eta_z, up_r, up_z);
      pgrad(um_xi, um_eta, xi_r, xi_z, eta_r, eta_z, um_r, um_z);
      pgrad(uz_xi, uz_eta, xi_r, xi_z, eta_r, eta_z, uz_r, uz_z);
      const float r = o.qr;
      const bool on_axis = r == 0.0f;
      const float up_over_r = on_axis ? up_r : U[k] / r;
      const float uz_over_r = on_axis ? uz_r : U[2 * NPE + k] / r;
      const float ur_r = (up_r + um_r) * is2, ur_z = (up_z + um_z) * is2;
      const float ut_r = (up_r - um_r) * is2, ut_z = (up_z - um_z) * is2;
      const float e_tt = s2 * up_over_r;
      const float lambda = o.k0, mu = o.k1;
      const float tr = ur_r + uz_z + e_tt;
      const float srr = lambda * tr + 2 * mu * ur_r;
      const float szz = lambda * tr + 2 * mu * uz_z;
      const float stt = lambda * tr + 2 * mu * e_tt;
      const float srz = mu * (ur_z + uz_r);
      const float srt = mu * (ut_r - s2 * up_over_r);
      const float stz = mu * (ut_z - uz_over_r);
      const float w = o.qw;
      float a0, a1;
      rflux(w, (srr + srt) * is2, (srz + stz) * is2, xi_r, xi_z, eta_r, eta_z, a0, a1);
      F[0 * NPE + k] = a0;
      F[1 * NPE + k] = a1;
      rflux(w, (srr - srt) * is2, (srz - stz) * is2, xi_r, xi_z, eta_r, eta_z, a0, a1);
      F[2 * NPE + k] = a0;
      F[3 * NPE + k] = a1;
      rflux(w, srz, szz, xi_r, xi_z, eta_r, eta_z, a0, a1);
      F[4 * NPE + k] = a0;
      F[5 * NPE + k] = a1;
      const float cp = s2 * (stt - srt), cz = -stz;
      if (axial && i == 0) {
        direct0 = direct1 = 0.0f;
        T.g[(le * NAX + 0) * NG + j] = w * cp;
        T.g[(le * NAX + 1) * NG + j] = w * cz;
      } else {
        const float hp = o.qh;
        direct0 = hp * cp;
        direct1 = hp * cz;
      }
    }
  }
}

// Weak divergence of the flux tile at point k (and the axis spread): the
// slot's force per component.
template <int NG, int PHYS>
__device__ __forceinline__ void element_out(const Tiles<NG, PHYS>& T, bool axial, int le, int k, int i, int j,
                                            const float* xc, const float* yc, float direct0, float direct1,
                                            float* fo) {
  constexpr int NPE = NG * NG;
  constexpr int C = Shape<PHYS>::C;
  constexpr int NAX = Shape<PHYS>::NAX;
  const float* F = T.f + le * 2 * C * NPE;
  if constexpr (PHYS == kElasticCart) {
    fo[0] = -wdiv<NG>(xc, yc, F, F + NPE, i, j);
    fo[1] = -wdiv<NG>(xc, yc, F + 2 * NPE, F + 3 * NPE, i, j);
  } else if constexpr (PHYS == kAcousticCart || PHYS == kAcousticAxi0) {
    fo[0] = -wdiv<NG>(xc, yc, F, F + NPE, i, j);
  } else if constexpr (PHYS == kAcousticAxi1) {
    float fe = -(wdiv<NG>(xc, yc, F, F + NPE, i, j) + direct0);
    if (axial) fe = spread<NG>(fe, i, j, T.g + le * NG, T.xr + le * NG, T.er + le * NG, T.Dj, T.D);
    fo[0] = fe;
  } else if constexpr (PHYS == kElasticAxi0) {
    float fr = -(wdiv<NG>(xc, yc, F, F + NPE, i, j) + direct0);
    if (axial) fr = spread<NG>(fr, i, j, T.g + le * NG, T.xr + le * NG, T.er + le * NG, T.Dj, T.D);
    fo[0] = fr;
    fo[1] = -wdiv<NG>(xc, yc, F + 2 * NPE, F + 3 * NPE, i, j);
  } else {
    float fp = -(wdiv<NG>(xc, yc, F, F + NPE, i, j) + direct0);
    const float fm = -wdiv<NG>(xc, yc, F + 2 * NPE, F + 3 * NPE, i, j);
    float fz = -(wdiv<NG>(xc, yc, F + 4 * NPE, F + 5 * NPE, i, j) + direct1);
    if (axial) {
      fp = spread<NG>(fp, i, j, T.g + (le * NAX + 0) * NG, T.xr + le * NG, T.er + le * NG, T.Dj, T.D);
      fz = spread<NG>(fz, i, j, T.g + (le * NAX + 1) * NG, T.xr + le * NG, T.er + le * NG, T.Dj, T.D);
    }
    fo[0] = fp;
    fo[1] = fm;
    fo[2] = fz;
  }
}

template <int NG, int PHYS>
__device__ __forceinline__ void stage_matrices(const ForceArgs& A, Tiles<NG, PHYS>& T) {
  constexpr int NPE = NG * NG;
  for (int q = threadIdx.x; q < NPE; q += kWG) {
    T.D[q] = A.D[q];
    T.Dj[q] = A.Dj[q];
    T.W2[q] = A.w2[q];
  }
}

// ------------------------------------------------------------------ seam

// Operator seam: element-local input in the reference layout, one batch per
// work-group, forces to f_local at elem[b] (out_ref) or out_pos0 + b.
template <int NG, int PHYS, bool RECT>
__global__ __launch_bounds__(kWG) void k_forces_local(ForceArgs A) {
  constexpr int NPE = NG * NG;
  constexpr int EPB = kWG / NPE;
  constexpr int C = Shape<PHYS>::C;
  __shared__ Tiles<NG, PHYS> T;
  const int t = threadIdx.x;
  const int le = t / NPE;
  const int k = t - le * NPE;
  const int i = k % NG, j = k / NG;
  const int b = xcd_contiguous(blockIdx.x, gridDim.x) * EPB + le;
  const bool act = le < EPB && b < A.nb;
  stage_matrices(A, T);
  Ops o;
  if (act) {
    load_ops<PHYS, RECT>(A, b, b * NPE + k, o);
    const size_t e = static_cast<size_t>(A.elem[b]);
#pragma unroll
    for (int c = 0; c < C; ++c) T.u[le * C * NPE + c * NPE + k] = A.ul[(e * C + c) * NPE + k];
  }
  __syncthreads();
  float xr[NG], yr[NG], xc[NG], yc[NG];
  d_rows(T, o.axial, i, j, xr, yr);
  if constexpr (!Shape<PHYS>::AXI) d_cols(T, o.axial, i, j, xc, yc);
  float direct0 = 0.0f, direct1 = 0.0f;
  if (act) element_flux<NG, PHYS, RECT>(T, o, le, k, i, j, xr, yr, direct0, direct1);
  __syncthreads();
  if (!act) return;
  if constexpr (Shape<PHYS>::AXI) d_cols(T, o.axial, i, j, xc, yc);
  float fo[C];
  element_out<NG, PHYS>(T, o.axial, le, k, i, j, xc, yc, direct0, direct1, fo);
  const size_t pos = A.out_ref ? static_cast<size_t>(A.elem[b]) : static_cast<size_t>(A.out_pos0) + b;
  float* out = A.fl + pos * C * NPE;
#pragma unroll
  for (int c = 0; c < C; ++c) out[c * NPE + k] = fo[c];
}

// ------------------------------------------------------------ production

// Registers of one element slot.
// (m, o, code, upd, uu, pv, pim are set for active slots only.)
template <int C>
struct Slot {
  bool act = false;
  int b = 0;
  int bm = -1;  // the batch's one model (batches of one model), uniform; else -1
  int32_t m;  // global point (< 0: constrained slot of row -1 - m)
  Ops o;
  int code;  // slot code (grouped / batch-local)
  bool upd;  // the slot updates its point (owned point or batch-local leader)
  float uu[C];
  float pv[C];
  float pim;
};

// First round trip: the slot's point, operands and rank.
template <int NG, int PHYS, bool RECT>
__device__ __forceinline__ void stage_point(const ForceArgs& A, const StepParams& P, int step, int batch, int le,
                                            int k, Slot<Shape<PHYS>::C>& s) {
  constexpr int NPE = NG * NG;
  constexpr int EPB = kWG / NPE;
  const int b = batch * EPB + le;
  s.act = le < EPB && b < A.nb;
  s.b = b;
  if (s.act) {
    const uint32_t slot = b * NPE + k;
    s.m = *at32(A.l2g, slot);
    load_ops<PHYS, RECT>(A, b, slot, s.o);
    s.code = *at32(A.code, slot);
  }
}

// Second round trip: the gathered field value and, for an owned point in a
// time step, its v and 1/m.  Lanes without a use read point 0 (no branch,
// so that the waits for these values do not also cover later loads).
template <int C>
__device__ __forceinline__ void stage_gather(const ForceArgs& A, int fuse, Slot<C>& s) {
  if (!s.act) return;
  const int32_t m = s.m;
  const uint32_t ng = static_cast<uint32_t>(A.ng);
  const uint32_t mg = m >= 0 ? m : 0;
#pragma unroll
  for (int c = 0; c < C; ++c) s.uu[c] = *at32(A.u, c * ng + mg);
  s.upd = fuse && m >= 0 && m < A.n_int && (!(s.code & kSlotLocal) || (s.code & kSlotLeader));
  const uint32_t mo = s.upd ? m : 0;
#pragma unroll
  for (int c = 0; c < C; ++c) s.pv[c] = *at32(A.v, c * ng + mo);
  s.pim = *at32(A.im, mo);
  if (m < 0) {
    const int32_t r = -1 - m;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      float v = 0.0f;
      for (int32_t q = A.row_off[r]; q < A.row_off[r + 1]; ++q) v += A.row_w[q] * A.u[c * A.ng + A.row_dofs[q]];
      s.uu[c] = v;
    }
  }
}

// Update of the slot's point with its assembled force f: a = f m^-1,
// corrector, energy terms, next predictor, following the tail's
// expressions in the same order (update.hip), with the model's own
// coefficients on a batch (whose last step predicts nothing).
template <int C, int NPE>
__device__ __forceinline__ void point_update(const ForceArgs& A, const StepParams& P, int step, const Slot<C>& s,
                                             int k, const float* U, const float* f, bool energy, double& ep,
                                             double& ek) {
  const uint32_t g = s.m;
  const uint32_t ng = static_cast<uint32_t>(A.ng);
  float rdt = P.rdt, hdt = P.hdt, hdt2 = P.hdt2;
  double cc = P.cc;
  int pred = P.pred;
  auto take = [&](const ModelCoef& mc) {
    rdt = mc.rdt;
    hdt = mc.hdt;
    hdt2 = mc.hdt2;
    cc = mc.cc;
    pred = pred && step < mc.n_steps;
  };
  if (A.models) {
    if (s.bm >= 0)
      take(A.models[s.bm]);  // uniform: scalar loads
    else
      take(A.models[A.emodel[s.b]]);
  }
  float u0[C], vc[C], ac[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint32_t o = c * ng + g;
    float u = U[c * NPE + k], v = s.pv[c];
    const float a = f[c] * s.pim;
    if (P.corr) v = v + hdt * a;
    u0[c] = u;
    vc[c] = v;
    ac[c] = a;
    if (pred) {
      u = u + (rdt * v + hdt2 * a);
      v = v + hdt * a;
      *at32(A.uw, o) = u;
    } else {
      *at32(A.a, o) = a;
    }
    *at32(A.v, o) = v;
  }
  // Energy terms (u.f with the step's u, kinetic with the corrected v), each
  // sum in component order; all under one uniform branch.
  if (energy) {
    const double m = *at32(A.mass, g);
#pragma unroll
    for (int c = 0; c < C; ++c) ep += static_cast<double>(u0[c]) * static_cast<double>(f[c]);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const double vd = vc[c], ad = ac[c];
      ek += m * (0.5 * vd * vd - cc * ad * ad);
    }
  }
}

// Stores the slot's force fo where its point's assembly reads it: a
// batch-local slot in the batch's buffer fob (for the point's leader), a
// grouped slot in its rank's group row, any other slot in f_local.  An owned
// point keeps its force for the update (point_update).
template <int C, int NPE>
__device__ __forceinline__ void slot_epilogue(const ForceArgs& A, const StepParams& P, const Slot<C>& s, int k,
                                              const float* fo, float* fob) {
  if (P.fuse && (s.code & kSlotLocal)) {
    const int z = local_place(s.code);
#pragma unroll
    for (int c = 0; c < C; ++c) fob[z * C + c] = fo[c];
    return;
  }
  if (s.upd) return;
  if (P.fuse && (s.code & kSlotGrouped)) {
    // The slot's place in the group rows: its rank's row, its point's column.
    const uint32_t q = static_cast<uint32_t>(s.m - A.gbegin) + static_cast<uint32_t>(s.code & 7) * A.gstride;
    const uint32_t cs = static_cast<uint32_t>(A.grows) * A.gstride;
#pragma unroll
    for (int c = 0; c < C; ++c) *at32(A.contrib, q + c * cs) = fo[c];
    return;
  }
  const uint32_t out = (static_cast<uint32_t>(A.out_pos0) + s.b) * (C * NPE) + k;
#pragma unroll
  for (int c = 0; c < C; ++c) *at32(A.fl, out + c * NPE) = fo[c];
}

// One batch of a block from the global field.
template <int NG, int PHYS, bool RECT>
__device__ __forceinline__ void force_batch(const ForceArgs& A, const StepParams& P, int step, Tiles<NG, PHYS>& T,
                                            int batch, bool energy, double& ep, double& ek) {
  constexpr int NPE = NG * NG;
  constexpr int C = Shape<PHYS>::C;
  // In a time step of a batch: the batch's last step and its one model
  // (-1: elements of several models), loaded as vector loads (the index
  // passes through a vector register) so that they are issued and waited
  // for with the slots' first loads rather than holding the work-group's
  // first scalar wait.  A batch whose models are all past their last step
  // leaves before it stores anything; a batch of one model (the usual case)
  // takes its model's coefficients as uniform values; a batch of several
  // models checks each element's model.
  const bool in_batch = P.fuse && A.batch_last && batch * (kWG / (NG * NG)) < A.nb;
  int last_v = 0, bm_v = -1;
  if (in_batch) {
    int bi;
    asm volatile("v_mov_b32 %0, %1" : "=v"(bi) : "s"(batch));
    last_v = A.batch_last[bi];
    bm_v = A.batch_model[bi];
  }
  const int t = threadIdx.x;
  const int le = t / NPE;
  const int k = t - le * NPE;
  const int i = k % NG, j = k / NG;
  stage_matrices(A, T);
  Slot<C> s;
  stage_point<NG, PHYS, RECT>(A, P, step, batch, le, k, s);
  if (in_batch) {
    if (step > __builtin_amdgcn_readfirstlane(last_v)) return;
    s.bm = __builtin_amdgcn_readfirstlane(bm_v);
    if (s.act && s.bm < 0) s.act = step <= A.models[A.emodel[s.b]].n_steps;
  }
  stage_gather<C>(A, P.fuse, s);
  float* U = T.u + le * C * NPE;
  if (s.act) {
#pragma unroll
    for (int c = 0; c < C; ++c) U[c * NPE + k] = s.uu[c];
  }
  __syncthreads();
  float xr[NG], yr[NG], xc[NG], yc[NG];
  d_rows(T, s.o.axial, i, j, xr, yr);
  if constexpr (!Shape<PHYS>::AXI) d_cols(T, s.o.axial, i, j, xc, yc);
  float direct0 = 0.0f, direct1 = 0.0f;
  if (s.act) element_flux<NG, PHYS, RECT>(T, s.o, le, k, i, j, xr, yr, direct0, direct1);
  __syncthreads();
  float fo[C];
  if (s.act) {
    if constexpr (Shape<PHYS>::AXI) d_cols(T, s.o.axial, i, j, xc, yc);
    element_out<NG, PHYS>(T, s.o.axial, le, k, i, j, xc, yc, direct0, direct1, fo);
    slot_epilogue<C, NPE>(A, P, s, k, fo, T.fo);
  }
  if (P.fuse) {
    // The assembled force of an updated point: its one slot's force (0.0f +
    // f, the assembly's sum of one term), or for a batch-local point the sum
    // from 0.0f of its slots' forces in slot order (the assembly's order),
    // taken by the point's leader once the batch has stored them.
    if (A.has_local) __syncthreads();
    if (s.act && s.upd) {
      float f[C];
#pragma unroll
      for (int c = 0; c < C; ++c) f[c] = 0.0f;
      if (s.code & kSlotLocal) {
        const int z = local_place(s.code), n = local_count(s.code);
        for (int r = 0; r < n; ++r) {
#pragma unroll
          for (int c = 0; c < C; ++c) f[c] += T.fo[(z + r) * C + c];
        }
      } else {
#pragma unroll
        for (int c = 0; c < C; ++c) f[c] += fo[c];
      }
      point_update<C, NPE>(A, P, step, s, k, U, f, energy, ep, ek);
    }
  }
}

// Fixed-tree work-group sums of the energy terms into out[0..1] (every
// work-item takes part).
__device__ __forceinline__ void write_energy(double ep, double ek, double* red, double* out) {
  const int t = threadIdx.x;
  const double v[2] = {ep, ek};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    __syncthreads();
    red[t] = v[q];
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) out[q] = red[0];
  }
}

// One block per launch.  In a time step the first launch also carries the
// previous step's receiver tasks in its first n_rec work-groups (they read
// the side buffers before this step's tails rewrite them).
// Waves per SIMD the gathering force kernel is compiled for: the eight
// that the hardware holds wherever the registers fit without spilling (the
// kernel waits on two dependent global round trips per work-group, so
// resident work-groups are what hides them).
constexpr int force_waves(int ng, int phys) {
  if (phys == kElasticAxi1) return ng <= 5 ? 8 : ng == 6 ? 7 : ng == 7 ? 6 : 5;
  if (phys == kElasticAxi0) return ng <= 6 ? 8 : ng <= 8 ? 7 : 6;
  return ng <= 7 ? 8 : ng == 8 ? 7 : 6;
}

template <int NG, int PHYS, bool RECT>
struct ForceLds {
  static constexpr size_t tiles = sizeof(Tiles<NG, PHYS>);
  static constexpr size_t rec = kWG * (sizeof(double) + sizeof(float));
  static constexpr size_t bytes = tiles > rec ? tiles : rec;
};

template <int NG, int PHYS, bool RECT>
__global__ __launch_bounds__(kWG, force_waves(NG, PHYS)) void k_forces(ForceArgs A, StepParams P, ReceiverDev R,
                                                                       int n_rec, double half_dt) {
  __shared__ __align__(16) unsigned char smem[ForceLds<NG, PHYS, RECT>::bytes];
  __shared__ double red[kWG];
  const int step = P.fuse ? clock_step(P.clk) : 0;
  const bool energy = P.fuse && clock_energy(P.clk, step);
  double ep = 0.0, ek = 0.0;
  const int wg = blockIdx.x;
  if (wg < n_rec) {
    double* sw = reinterpret_cast<double*>(smem);
    float* sx = reinterpret_cast<float*>(smem + kWG * sizeof(double));
    receiver_task<kWG>(R, wg, P.clk, step - 1, 1, half_dt, sw, sx);
  } else {
    force_batch<NG, PHYS, RECT>(A, P, step, *reinterpret_cast<Tiles<NG, PHYS>*>(smem),
                                xcd_contiguous(wg - n_rec, gridDim.x - n_rec), energy, ep, ek);
  }
  if (energy) write_energy(ep, ek, red, P.partial + 2 * blockIdx.x);
}

// The elastic block A and the acoustic block B of a time step in one
// launch (the two are independent: each reads the predicted field of its
// own domain): the work-groups after the receiver tasks ([0, n_rec), padded
// to rec_end, a multiple of 8) take A's batches ([rec_end, rec_end + na8),
// na8 a multiple of 8), then B's, each part in the XCD-contiguous order.
// B's smaller work-groups then fill the end of A's instead of running as a
// launch of their own.
constexpr int pair_waves(int ng, int pa, int pb) {
  return force_waves(ng, pa) < force_waves(ng, pb) ? force_waves(ng, pa) : force_waves(ng, pb);
}

// WAVES: the waves per SIMD the kernel's registers are budgeted for (the
// occupancy variants of the autotuner; pair_waves by default).
template <int NG, int PA, bool RA, int PB, bool RB, int WAVES = pair_waves(NG, PA, PB)>
__global__ __launch_bounds__(kWG, WAVES) void k_forces_pair(ForceArgs A, ForceArgs B, int na8, StepParams P,
                                                          ReceiverDev R, int n_rec, int rec_end, double half_dt) {
  constexpr size_t la = ForceLds<NG, PA, RA>::bytes, lb = ForceLds<NG, PB, RB>::bytes;
  __shared__ __align__(16) unsigned char smem[la > lb ? la : lb];
  __shared__ double red[kWG];
  const int step = P.fuse ? clock_step(P.clk) : 0;
  const bool energy = P.fuse && clock_energy(P.clk, step);
  double ep = 0.0, ek = 0.0;
  const int wg = blockIdx.x;
  if (wg < rec_end) {
    if (wg < n_rec) {
      double* sw = reinterpret_cast<double*>(smem);
      float* sx = reinterpret_cast<float*>(smem + kWG * sizeof(double));
      receiver_task<kWG>(R, wg, P.clk, step - 1, 1, half_dt, sw, sx);
    }
  } else if (wg - rec_end < na8) {
    force_batch<NG, PA, RA>(A, P, step, *reinterpret_cast<Tiles<NG, PA>*>(smem), xcd_contiguous(wg - rec_end, na8),
                            energy, ep, ek);
  } else {
    const int nb8 = gridDim.x - rec_end - na8;
    force_batch<NG, PB, RB>(B, P, step, *reinterpret_cast<Tiles<NG, PB>*>(smem),
                            xcd_contiguous(wg - rec_end - na8, nb8), energy, ep, ek);
  }
  if (energy) write_energy(ep, ek, red, P.partial + 2 * blockIdx.x);
}

// Extra work of a force launch: receiver tasks in its first work-groups.
struct RecWork {
  const ReceiverDev* r = nullptr;
  int n = 0;
  double half_dt = 0.0;
};

template <int NG, int PHYS, bool RECT, bool GATHER>
hipError_t launch_one(const ForceArgs& a, const StepParams& p, const RecWork& rw, hipStream_t s, int* grid_out) {
  constexpr int EPB = kWG / (NG * NG);
  const int batches = (a.nb + EPB - 1) / EPB;
  if constexpr (GATHER) {
    const int n_rec = rw.r ? rw.n : 0;
    const int grid = batches + n_rec;
    if (grid_out) *grid_out = grid;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_forces<NG, PHYS, RECT>), dim3(grid), dim3(kWG), 0, s, a, p,
                       rw.r ? *rw.r : ReceiverDev(), n_rec, rw.half_dt);
  } else {
    if (grid_out) *grid_out = batches;
    if (batches == 0) return hipSuccess;
    hipLaunchKernelGGL((k_forces_local<NG, PHYS, RECT>), dim3(batches), dim3(kWG), 0, s, a);
  }
  return hipGetLastError();
}

template <int NG, bool GATHER>
hipError_t launch_ng(int phys, bool rect, const ForceArgs& a, const StepParams& p, const RecWork& w, hipStream_t s,
                     int* g) {
  switch (phys) {
    case kAcousticCart:
      return rect ? launch_one<NG, kAcousticCart, true, GATHER>(a, p, w, s, g)
                  : launch_one<NG, kAcousticCart, false, GATHER>(a, p, w, s, g);
    case kElasticCart:
      return rect ? launch_one<NG, kElasticCart, true, GATHER>(a, p, w, s, g)
                  : launch_one<NG, kElasticCart, false, GATHER>(a, p, w, s, g);
    case kAcousticAxi0: return launch_one<NG, kAcousticAxi0, false, GATHER>(a, p, w, s, g);
    case kElasticAxi0: return launch_one<NG, kElasticAxi0, false, GATHER>(a, p, w, s, g);
    case kAcousticAxi1: return launch_one<NG, kAcousticAxi1, false, GATHER>(a, p, w, s, g);
    case kElasticAxi1: return launch_one<NG, kElasticAxi1, false, GATHER>(a, p, w, s, g);
  }
  return hipErrorInvalidValue;
}

template <bool GATHER>
hipError_t launch_gather_mode(int ngll, int phys, bool rect, const ForceArgs& a, const StepParams& p, const RecWork& w,
                              hipStream_t s, int* g) {
  switch (ngll) {
    case 2: return launch_ng<2, GATHER>(phys, rect, a, p, w, s, g);
    case 3: return launch_ng<3, GATHER>(phys, rect, a, p, w, s, g);
    case 4: return launch_ng<4, GATHER>(phys, rect, a, p, w, s, g);
    case 5: return launch_ng<5, GATHER>(phys, rect, a, p, w, s, g);
    case 6: return launch_ng<6, GATHER>(phys, rect, a, p, w, s, g);
    case 7: return launch_ng<7, GATHER>(phys, rect, a, p, w, s, g);
    case 8: return launch_ng<8, GATHER>(phys, rect, a, p, w, s, g);
    case 9: return launch_ng<9, GATHER>(phys, rect, a, p, w, s, g);
  }
  return hipErrorInvalidValue;
}

}  // namespace

namespace {

// The LDS a work-group of the paired kernel declares.
template <int NG, int PA, bool RA, int PB, bool RB>
constexpr size_t pair_lds() {
  constexpr size_t la = ForceLds<NG, PA, RA>::bytes, lb = ForceLds<NG, PB, RB>::bytes;
  return (la > lb ? la : lb) + kWG * sizeof(double);
}

// wgs_per_cu > 0: the launch asks for enough dynamic LDS (unused) that at
// most that many work-groups, i.e. waves per SIMD, fit on a CU's 160 KiB.
template <int NG, int PA, bool RA, int PB, bool RB, int WAVES = pair_waves(NG, PA, PB)>
hipError_t launch_pair_one(const ForceArgs& a, const ForceArgs& b, const StepParams& p, const RecWork& rw,
                           hipStream_t s, int* grid_out, int wgs_per_cu = 0) {
  constexpr int EPB = kWG / (NG * NG);
  auto up8 = [](int n) { return (n + 7) / 8 * 8; };
  const int ba = (a.nb + EPB - 1) / EPB, bb = (b.nb + EPB - 1) / EPB;
  const int n_rec = rw.r ? rw.n : 0;
  const int rec_end = up8(n_rec), na8 = up8(ba);
  const int grid = rec_end + na8 + bb;
  if (grid_out) *grid_out = grid;
  if (grid == 0) return hipSuccess;
  size_t dyn = 0;
  if (wgs_per_cu > 0) {
    constexpr size_t kLdsPerCu = 160 * 1024, st = pair_lds<NG, PA, RA, PB, RB>();
    const size_t want = (kLdsPerCu / (wgs_per_cu + 1) + 1024) / 512 * 512;  // > total / (n + 1)
    dyn = want > st ? want - st : 0;
  }
  hipLaunchKernelGGL((k_forces_pair<NG, PA, RA, PB, RB, WAVES>), dim3(grid), dim3(kWG), dyn, s, a, b, na8, p,
                     rw.r ? *rw.r : ReceiverDev(), n_rec, rec_end, rw.half_dt);
  return hipGetLastError();
}

template <int NG>
hipError_t launch_pair_ng(int pa, bool ra, int pb, bool rb, const ForceArgs& a, const ForceArgs& b,
                          const StepParams& p, const RecWork& w, hipStream_t s, int* g) {
  if (pa == kElasticCart && pb == kAcousticCart) {
    if (ra && rb) return launch_pair_one<NG, kElasticCart, true, kAcousticCart, true>(a, b, p, w, s, g);
    if (ra) return launch_pair_one<NG, kElasticCart, true, kAcousticCart, false>(a, b, p, w, s, g);
    if (rb) return launch_pair_one<NG, kElasticCart, false, kAcousticCart, true>(a, b, p, w, s, g);
    return launch_pair_one<NG, kElasticCart, false, kAcousticCart, false>(a, b, p, w, s, g);
  }
  if (pa == kElasticAxi0 && pb == kAcousticAxi0)
    return launch_pair_one<NG, kElasticAxi0, false, kAcousticAxi0, false>(a, b, p, w, s, g);
  if (pa == kElasticAxi1 && pb == kAcousticAxi1)
    return launch_pair_one<NG, kElasticAxi1, false, kAcousticAxi1, false>(a, b, p, w, s, g);
  return hipErrorInvalidValue;
}


// Occupancy variants of the paired force launch for the autotuner (orders 4
// to 6): level 0 is the default launch (registers budgeted for pair_waves
// waves per SIMD, as many work-groups per CU as fit); level 1 budgets the
// registers for at most 6 waves per SIMD and limits the launch to 6
// work-groups per CU, level 2 to 4 (through unused dynamic LDS).  Fewer
// resident work-groups leave each more of the L2 and more registers.  This
// file is compiled twice, with
// LLVM's max-ilp scheduling strategy (the library's default, which defines
// launch_forces_pair_max_ilp and everything else) and with the default
// scheduler (forces_default_sched.hip: the variants only).
constexpr int occ_waves(int level, int def) {
  return level == 0 ? def : level == 1 ? (def < 6 ? def : 6) : (def < 4 ? def : 4);
}

template <int NG, int PA, bool RA, int PB, bool RB>
hipError_t launch_pair_occ(int level, const ForceArgs& a, const ForceArgs& b, const StepParams& p, const RecWork& w,
                           hipStream_t s, int* g) {
  constexpr int d = pair_waves(NG, PA, PB);
  switch (level) {
    case 0: return launch_pair_one<NG, PA, RA, PB, RB, occ_waves(0, d)>(a, b, p, w, s, g);
    case 1: return launch_pair_one<NG, PA, RA, PB, RB, occ_waves(1, d)>(a, b, p, w, s, g, occ_waves(1, d));
    case 2: return launch_pair_one<NG, PA, RA, PB, RB, occ_waves(2, d)>(a, b, p, w, s, g, occ_waves(2, d));
  }
  return hipErrorInvalidValue;
}

template <int NG>
hipError_t launch_pair_tuned_ng(int level, int pa, bool ra, int pb, bool rb, const ForceArgs& a, const ForceArgs& b,
                                const StepParams& p, const RecWork& w, hipStream_t s, int* g) {
  if (pa == kElasticCart && pb == kAcousticCart) {
    if (ra && rb) return launch_pair_occ<NG, kElasticCart, true, kAcousticCart, true>(level, a, b, p, w, s, g);
    if (ra) return launch_pair_occ<NG, kElasticCart, true, kAcousticCart, false>(level, a, b, p, w, s, g);
    if (rb) return launch_pair_occ<NG, kElasticCart, false, kAcousticCart, true>(level, a, b, p, w, s, g);
    return launch_pair_occ<NG, kElasticCart, false, kAcousticCart, false>(level, a, b, p, w, s, g);
  }
  if (pa == kElasticAxi0 && pb == kAcousticAxi0)
    return launch_pair_occ<NG, kElasticAxi0, false, kAcousticAxi0, false>(level, a, b, p, w, s, g);
  if (pa == kElasticAxi1 && pb == kAcousticAxi1)
    return launch_pair_occ<NG, kElasticAxi1, false, kAcousticAxi1, false>(level, a, b, p, w, s, g);
  return hipErrorInvalidValue;
}

}  // namespace

#ifndef INSEM_FORCES_TUNED_LAUNCH
#define INSEM_FORCES_TUNED_LAUNCH launch_forces_pair_max_ilp
#endif

hipError_t INSEM_FORCES_TUNED_LAUNCH(int ngll, int level, int pa, bool ra, int pb, bool rb, const ForceArgs& a,
                                     const ForceArgs& b, const StepParams& p, hipStream_t s, int* grid,
                                     const ReceiverDev* rec, int n_rec, double half_dt) {
  RecWork w;
  w.r = rec;
  w.n = n_rec;
  w.half_dt = half_dt;
  switch (ngll) {
    case 5: return launch_pair_tuned_ng<5>(level, pa, ra, pb, rb, a, b, p, w, s, grid);
    case 6: return launch_pair_tuned_ng<6>(level, pa, ra, pb, rb, a, b, p, w, s, grid);
    case 7: return launch_pair_tuned_ng<7>(level, pa, ra, pb, rb, a, b, p, w, s, grid);
  }
  return hipErrorInvalidValue;
}

#ifndef INSEM_FORCES_VARIANTS_ONLY

bool forces_pair_supported(int pa, bool ra, int pb, bool rb) {
  const bool cart = pa == kElasticCart && pb == kAcousticCart;
  const bool axi = (pa == kElasticAxi0 && pb == kAcousticAxi0) || (pa == kElasticAxi1 && pb == kAcousticAxi1);
  return cart || (axi && !ra && !rb);
}

namespace {

// The autotuner's table: the winning variant of each tuned paired launch
// (written by in-house-sem_amd/tune/autotune.py; a missing entry is variant 0).
// chunk: the steps per captured hipGraph chunk (0: the default).
struct TunedForce {
  int ngll, pa, ra, pb, rb, level, sched, chunk;
};
constexpr TunedForce kTunedForces[] = {
#include "tuned_forces.inc"
    {0, -1, 0, -1, 0, 0, 0, 0},  // end
};

const TunedForce* tuned_entry(int ngll, int pa, bool ra, int pb, bool rb) {
  for (const TunedForce& t : kTunedForces)
    if (t.ngll == ngll && t.pa == pa && t.ra == int(ra) && t.pb == pb && t.rb == int(rb)) return &t;
  return nullptr;
}

}  // namespace

int tuned_force_variant(int ngll, int pa, bool ra, int pb, bool rb) {
  const TunedForce* t = tuned_entry(ngll, pa, ra, pb, rb);
  return t ? t->level + kForceLevels * t->sched : 0;
}

int tuned_chunk_steps(int ngll, int pa, bool ra, int pb, bool rb) {
  const TunedForce* t = tuned_entry(ngll, pa, ra, pb, rb);
  return t ? t->chunk : 0;
}

hipError_t launch_forces_pair(int ngll, int pa, bool ra, int pb, bool rb, const ForceArgs& a, const ForceArgs& b,
                              const StepParams& p, hipStream_t s, int* grid, const ReceiverDev* rec, int n_rec,
                              double half_dt, int variant) {
  if (!p.fuse || !forces_pair_supported(pa, ra, pb, rb)) return hipErrorInvalidValue;
  if (variant < 0) variant = tuned_force_variant(ngll, pa, ra, pb, rb);
  if (variant < 0 || variant >= kForceVariants) return hipErrorInvalidValue;
  if (variant != 0 && ngll >= 5 && ngll <= 7) {
    const int level = variant % kForceLevels, sched = variant / kForceLevels;
    return sched == 0 ? launch_forces_pair_max_ilp(ngll, level, pa, ra, pb, rb, a, b, p, s, grid, rec, n_rec, half_dt)
                      : launch_forces_pair_default_sched(ngll, level, pa, ra, pb, rb, a, b, p, s, grid, rec, n_rec,
                                                         half_dt);
  }
  RecWork w;
  w.r = rec;
  w.n = n_rec;
  w.half_dt = half_dt;
  switch (ngll) {
    case 2: return launch_pair_ng<2>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 3: return launch_pair_ng<3>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 4: return launch_pair_ng<4>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 5: return launch_pair_ng<5>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 6: return launch_pair_ng<6>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 7: return launch_pair_ng<7>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 8: return launch_pair_ng<8>(pa, ra, pb, rb, a, b, p, w, s, grid);
    case 9: return launch_pair_ng<9>(pa, ra, pb, rb, a, b, p, w, s, grid);
  }
  return hipErrorInvalidValue;
}

int force_batch_elements(int ngll) { return kWG / (ngll * ngll); }

int force_blocks(int ngll, int nb) {
  const int epb = kWG / (ngll * ngll);
  return (nb + epb - 1) / epb;
}

hipError_t launch_forces(int ngll, int phys, bool rect, bool gather, const ForceArgs& a, const StepParams& p,
                         hipStream_t s, int* grid, const ReceiverDev* rec, int n_rec, double half_dt) {
  if (p.fuse && !gather) return hipErrorInvalidValue;
  RecWork w;
  w.r = rec;
  w.n = n_rec;
  w.half_dt = half_dt;
  return gather ? launch_gather_mode<true>(ngll, phys, rect, a, p, w, s, grid)
                : launch_gather_mode<false>(ngll, phys, rect, a, p, w, s, grid);
}

#endif  // INSEM_FORCES_VARIANTS_ONLY

}  // namespace insem_hip
