// Element internal-force kernels for gfx950: f_local = -K u_local of one
// element block, with the gather from the global field fused in.
//
// Mapping: one work-item per element point, 256 work-items per work-group
// holding floor(256 / npe) whole elements (10 at order 4).  The gathered
// element fields and the reference fluxes are staged in LDS; each work-item
// keeps the rows and columns of the derivative matrices it needs in
// registers.  Every per-point arithmetic expression and every summation
// follows the reference kernels (cpp/src/cpu/kernels.cpp:19-431) in the same
// order, and the library is compiled with -ffp-contract=off, so the result
// is bitwise the reference's float result.
//
// Work-group ids are remapped so that each of the 8 XCDs processes one
// contiguous range of elements (neighbouring elements share global points,
// which then stay in that XCD's L2).

#include "device.hpp"
#include "launch.hpp"

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

template <int NG, int PHYS, bool RECT, bool GATHER>
__global__ __launch_bounds__(kWG) void k_forces(ForceArgs A) {
  constexpr int NPE = NG * NG;
  constexpr int EPB = kWG / NPE;
  constexpr int C = phys_ncomp(PHYS);
  constexpr bool AXI = PHYS >= kAcousticAxi0;
  constexpr int NAX = PHYS == kElasticAxi1 ? 2 : 1;
  __shared__ float sD[NPE], sDj[NPE], sW2[NPE];
  __shared__ float su[EPB * C * NPE];
  __shared__ float sf[EPB * 2 * C * NPE];
  __shared__ float sg[AXI ? EPB * NAX * NG : 1];
  __shared__ float sxr[AXI ? EPB * NG : 1];
  __shared__ float ser[AXI ? EPB * NG : 1];

  const int t = threadIdx.x;
  const int le = t / NPE;
  const int k = t - le * NPE;
  const int i = k % NG, j = k / NG;
  const int bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int b = bid * EPB + le;
  const bool act = le < EPB && b < A.nb;

  // Global loads are issued in the order their results are needed (the
  // shared matrices, the slot's point, the per-point operands, then
  // everything that depends on the point), so that each wait covers only
  // what it must.
  for (int q = t; q < NPE; q += kWG) {
    sD[q] = A.D[q];
    sDj[q] = A.Dj[q];
    sW2[q] = A.w2[q];
  }
  int32_t slot_point = -1;  // global point of this slot (< 0: constrained)
  float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f, q3 = 0.0f, q4 = 0.0f, k0 = 0.0f, k1 = 0.0f;
  float qw = 0.0f, qh = 0.0f, qr = 0.0f;
  int prk = 0;              // fused step: rank of a grouped point's slot
  bool owned = false;       // fused step: this work-item updates the point
  float pv[C], pim = 0.0f;  // its v and 1/m, loaded with the gather
  float* U = su + le * C * NPE;
  float* F = sf + le * 2 * C * NPE;
  if (act) {
    int32_t m = 0;
    if constexpr (GATHER) m = A.l2g[static_cast<size_t>(b) * NPE + k];
    q0 = ld(A.m[0], b, k);
    q3 = ld(A.m[3], b, k);
    if constexpr (!AXI) q4 = ld(A.m[4], b, k);
    if constexpr (!RECT) {
      q1 = ld(A.m[1], b, k);
      q2 = ld(A.m[2], b, k);
    }
    k0 = ld(A.c0, b, k);
    if constexpr (C >= 2) k1 = ld(A.c1, b, k);
    if constexpr (AXI) {
      qw = ld(A.wr, b, k);
      qr = ld(A.r, b, k);
      if constexpr (PHYS != kAcousticAxi0) qh = ld(A.hoop, b, k);
    }
    float uu[C];
    if constexpr (GATHER) {
      prk = A.rank[static_cast<size_t>(b) * NPE + k];
      slot_point = m;
      // Branch-free loads (lanes without a use read point 0), so that the
      // waits for the gathered values do not also wait for later loads.
      const int64_t mg = m >= 0 ? m : 0;
#pragma unroll
      for (int c = 0; c < C; ++c) uu[c] = A.u[c * A.ng + mg];
      owned = A.fuse && m >= 0 && m < A.n_int;
      const int64_t mo = owned ? m : 0;
#pragma unroll
      for (int c = 0; c < C; ++c) pv[c] = A.v[c * A.ng + mo];
      pim = A.im[mo];
      if (m < 0) {
        const int32_t r = -1 - m;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          float s = 0.0f;
          for (int32_t q = A.row_off[r]; q < A.row_off[r + 1]; ++q) s += A.row_w[q] * A.u[c * A.ng + A.row_dofs[q]];
          uu[c] = s;
        }
      }
    } else {
      const size_t e = static_cast<size_t>(A.elem[b]);
#pragma unroll
      for (int c = 0; c < C; ++c) uu[c] = A.ul[(e * C + c) * NPE + k];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) U[c * NPE + k] = uu[c];
  }
  __syncthreads();

  // Derivative-matrix rows and columns of this point.
  bool axial = false;
  if constexpr (AXI) axial = act && A.axial[b] != 0;
  const float* Dx = axial ? sDj : sD;
  float xr[NG], yr[NG], xc[NG], yc[NG];
#pragma unroll
  for (int l = 0; l < NG; ++l) {
    xr[l] = Dx[i * NG + l];
    yr[l] = sD[j * NG + l];
    xc[l] = Dx[l * NG + i];
    yc[l] = sD[l * NG + j];
  }

  float direct0 = 0.0f, direct1 = 0.0f;
  if (act) {
    if constexpr (PHYS == kElasticCart) {
      float ux_xi, ux_eta, uz_xi, uz_eta;
      deriv<NG>(xr, yr, U, i, j, ux_xi, ux_eta);
      deriv<NG>(xr, yr, U + NPE, i, j, uz_xi, uz_eta);
      float xi_x, xi_z = 0.0f, eta_x = 0.0f, eta_z, jac;
      xi_x = q0;
      eta_z = q3;
      jac = q4;
      if constexpr (!RECT) {
        xi_z = q1;
        eta_x = q2;
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
      const float lambda = k0, mu = k1;
      const float div = ux_x + uz_z;
      const float sxx = lambda * div + 2.0f * mu * ux_x;
      const float szz = lambda * div + 2.0f * mu * uz_z;
      const float sxz = mu * (ux_z + uz_x);
      const float wj = sW2[k] * jac;
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
      xi_x = q0;
      eta_z = q3;
      jac = q4;
      if constexpr (!RECT) {
        xi_z = q1;
        eta_x = q2;
      }
      const float c = sW2[k] * jac * k0;
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
      const float xi_r = q0, xi_z = q1;
      const float eta_r = q2, eta_z = q3;
      if (axial && i == 0) {
        sxr[le * NG + j] = xi_r;
        ser[le * NG + j] = eta_r;
      }
      if constexpr (PHYS == kAcousticAxi0) {
        float p_xi, p_eta, p_r, p_z;
        deriv<NG>(xr, yr, U, i, j, p_xi, p_eta);
        pgrad(p_xi, p_eta, xi_r, xi_z, eta_r, eta_z, p_r, p_z);
        float f1, f2;
        rflux(qw * k0, p_r, p_z, xi_r, xi_z, eta_r, eta_z, f1, f2);
        F[k] = f1;
        F[NPE + k] = f2;
      } else if constexpr (PHYS == kAcousticAxi1) {
        float p_xi, p_eta, p_r, p_z;
        deriv<NG>(xr, yr, U, i, j, p_xi, p_eta);
        pgrad(p_xi, p_eta, xi_r, xi_z, eta_r, eta_z, p_r, p_z);
        const float inv_rho = k0;
        const float w = qw;
        float f1, f2;
        rflux(w * inv_rho, p_r, p_z, xi_r, xi_z, eta_r, eta_z, f1, f2);
        F[k] = f1;
        F[NPE + k] = f2;
        if (axial && i == 0) {
          direct0 = 0.0f;
          sg[le * NG + j] = w * inv_rho * p_r;
        } else {
          direct0 = qh * inv_rho * U[k] / qr;
        }
      } else if constexpr (PHYS == kElasticAxi0) {
        float ur_xi, ur_eta, uz_xi, uz_eta;
        deriv<NG>(xr, yr, U, i, j, ur_xi, ur_eta);
        deriv<NG>(xr, yr, U + NPE, i, j, uz_xi, uz_eta);
        float ur_r, ur_z, uz_r, uz_z;
        pgrad(ur_xi, ur_eta, xi_r, xi_z, eta_r, eta_z, ur_r, ur_z);
        pgrad(uz_xi, uz_eta, xi_r, xi_z, eta_r, eta_z, uz_r, uz_z);
        const float r = qr;
        const bool on_axis = r == 0.0f;
        const float e_tt = on_axis ? ur_r : U[k] / r;
        const float lambda = k0, mu = k1;
        const float tr = ur_r + uz_z + e_tt;
        const float srr = lambda * tr + 2 * mu * ur_r;
        const float szz = lambda * tr + 2 * mu * uz_z;
        const float stt = lambda * tr + 2 * mu * e_tt;
        const float srz = mu * (ur_z + uz_r);
        const float w = qw;
        float f1r, f2r, f1z, f2z;
        rflux(w, srr, srz, xi_r, xi_z, eta_r, eta_z, f1r, f2r);
        rflux(w, srz, szz, xi_r, xi_z, eta_r, eta_z, f1z, f2z);
        F[0 * NPE + k] = f1r;
        F[1 * NPE + k] = f2r;
        F[2 * NPE + k] = f1z;
        F[3 * NPE + k] = f2z;
        if (axial && i == 0) {
          direct0 = 0.0f;
          sg[le * NG + j] = w * stt;
        } else {
          direct0 = qh * stt;
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
        const float r = qr;
        const bool on_axis = r == 0.0f;
        const float up_over_r = on_axis ? up_r : U[k] / r;
        const float uz_over_r = on_axis ? uz_r : U[2 * NPE + k] / r;
        const float ur_r = (up_r + um_r) * is2, ur_z = (up_z + um_z) * is2;
        const float ut_r = (up_r - um_r) * is2, ut_z = (up_z - um_z) * is2;
        const float e_tt = s2 * up_over_r;
        const float lambda = k0, mu = k1;
        const float tr = ur_r + uz_z + e_tt;
        const float srr = lambda * tr + 2 * mu * ur_r;
        const float szz = lambda * tr + 2 * mu * uz_z;
        const float stt = lambda * tr + 2 * mu * e_tt;
        const float srz = mu * (ur_z + uz_r);
        const float srt = mu * (ut_r - s2 * up_over_r);
        const float stz = mu * (ut_z - uz_over_r);
        const float w = qw;
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
          sg[(le * NAX + 0) * NG + j] = w * cp;
          sg[(le * NAX + 1) * NG + j] = w * cz;
        } else {
          const float hp = qh;
          direct0 = hp * cp;
          direct1 = hp * cz;
        }
      }
    }
  }
  __syncthreads();

  float fo[C];
  if (act) {
    if constexpr (PHYS == kElasticCart) {
      fo[0] = -wdiv<NG>(xc, yc, F, F + NPE, i, j);
      fo[1] = -wdiv<NG>(xc, yc, F + 2 * NPE, F + 3 * NPE, i, j);
    } else if constexpr (PHYS == kAcousticCart || PHYS == kAcousticAxi0) {
      fo[0] = -wdiv<NG>(xc, yc, F, F + NPE, i, j);
    } else if constexpr (PHYS == kAcousticAxi1) {
      float fe = -(wdiv<NG>(xc, yc, F, F + NPE, i, j) + direct0);
      if (axial) fe = spread<NG>(fe, i, j, sg + le * NG, sxr + le * NG, ser + le * NG, sDj, sD);
      fo[0] = fe;
    } else if constexpr (PHYS == kElasticAxi0) {
      float fr = -(wdiv<NG>(xc, yc, F, F + NPE, i, j) + direct0);
      if (axial) fr = spread<NG>(fr, i, j, sg + le * NG, sxr + le * NG, ser + le * NG, sDj, sD);
      fo[0] = fr;
      fo[1] = -wdiv<NG>(xc, yc, F + 2 * NPE, F + 3 * NPE, i, j);
    } else {
      float fp = -(wdiv<NG>(xc, yc, F, F + NPE, i, j) + direct0);
      const float fm = -wdiv<NG>(xc, yc, F + 2 * NPE, F + 3 * NPE, i, j);
      float fz = -(wdiv<NG>(xc, yc, F + 4 * NPE, F + 5 * NPE, i, j) + direct1);
      if (axial) {
        fp = spread<NG>(fp, i, j, sg + (le * NAX + 0) * NG, sxr + le * NG, ser + le * NG, sDj, sD);
        fz = spread<NG>(fz, i, j, sg + (le * NAX + 1) * NG, sxr + le * NG, ser + le * NG, sDj, sD);
      }
      fo[0] = fp;
      fo[1] = fm;
      fo[2] = fz;
    }
  }

  bool done = false;
  if constexpr (GATHER) {
    if (A.fuse) {
      const bool energy = clock_energy(A.clk, clock_step(A.clk));
      double ep = 0.0, ek = 0.0;
      if (act && owned) {
        // The slot's force is the point's assembled force (0.0f + f, the
        // assembly's sum of one term); the update follows the tail's
        // expressions in the same order (update.hip).
        done = true;
        const int64_t g = slot_point;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int64_t o = c * A.ng + g;
          float u = U[c * NPE + k], v = pv[c];
          const float f = 0.0f + fo[c];
          const float a = f * pim;
          if (energy) ep += static_cast<double>(u) * static_cast<double>(f);
          if (A.corr) v = v + A.hdt * a;
          if (energy) {
            const double vd = v, ad = a;
            ek += static_cast<double>(A.mass[g]) * (0.5 * vd * vd - A.cc * ad * ad);
          }
          if (A.pred) {
            u = u + (A.rdt * v + A.hdt2 * a);
            v = v + A.hdt * a;
            A.uw[o] = u;
          } else {
            A.a[o] = a;
          }
          A.v[o] = v;
        }
      } else if (act && slot_point >= A.grp.begin && slot_point < A.grp.end) {
        done = true;
        int64_t off, n;
        int cnt;
        group_locate(A.grp, slot_point, &off, &n, &cnt);
        float* dst = A.contrib + off + prk * n;
#pragma unroll
        for (int c = 0; c < C; ++c) dst[c * cnt * n] = fo[c];
      }
      if (energy) {
        // Fixed-tree work-group sums (every work-item takes part).
        __shared__ double red[kWG];
        double s2[2] = {ep, ek};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          red[t] = s2[q];
          __syncthreads();
          for (int w = kWG / 2; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
          }
          if (t == 0) A.partial[2 * blockIdx.x + q] = red[0];
          __syncthreads();
        }
      }
    }
  }
  if (!act || done) return;

  const size_t pos = A.out_ref ? static_cast<size_t>(A.elem[b]) : static_cast<size_t>(A.out_pos0) + b;
  float* out = A.fl + pos * C * NPE;
#pragma unroll
  for (int c = 0; c < C; ++c) out[c * NPE + k] = fo[c];
}

template <int NG, int PHYS, bool RECT, bool GATHER>
hipError_t launch_one(const ForceArgs& a, hipStream_t s) {
  constexpr int EPB = kWG / (NG * NG);
  const int grid = (a.nb + EPB - 1) / EPB;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_forces<NG, PHYS, RECT, GATHER>), dim3(grid), dim3(kWG), 0, s, a);
  return hipGetLastError();
}

template <int NG, bool GATHER>
hipError_t launch_ng(int phys, bool rect, const ForceArgs& a, hipStream_t s) {
  switch (phys) {
    case kAcousticCart:
      return rect ? launch_one<NG, kAcousticCart, true, GATHER>(a, s) : launch_one<NG, kAcousticCart, false, GATHER>(a, s);
    case kElasticCart:
      return rect ? launch_one<NG, kElasticCart, true, GATHER>(a, s) : launch_one<NG, kElasticCart, false, GATHER>(a, s);
    case kAcousticAxi0: return launch_one<NG, kAcousticAxi0, false, GATHER>(a, s);
    case kElasticAxi0: return launch_one<NG, kElasticAxi0, false, GATHER>(a, s);
    case kAcousticAxi1: return launch_one<NG, kAcousticAxi1, false, GATHER>(a, s);
    case kElasticAxi1: return launch_one<NG, kElasticAxi1, false, GATHER>(a, s);
  }
  return hipErrorInvalidValue;
}

template <bool GATHER>
hipError_t launch_gather_mode(int ngll, int phys, bool rect, const ForceArgs& a, hipStream_t s) {
  switch (ngll) {
    case 2: return launch_ng<2, GATHER>(phys, rect, a, s);
    case 3: return launch_ng<3, GATHER>(phys, rect, a, s);
    case 4: return launch_ng<4, GATHER>(phys, rect, a, s);
    case 5: return launch_ng<5, GATHER>(phys, rect, a, s);
    case 6: return launch_ng<6, GATHER>(phys, rect, a, s);
    case 7: return launch_ng<7, GATHER>(phys, rect, a, s);
    case 8: return launch_ng<8, GATHER>(phys, rect, a, s);
    case 9: return launch_ng<9, GATHER>(phys, rect, a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_forces(int ngll, int phys, bool rect, bool gather, const ForceArgs& a, hipStream_t s) {
  if (a.fuse && !gather) return hipErrorInvalidValue;
  return gather ? launch_gather_mode<true>(ngll, phys, rect, a, s) : launch_gather_mode<false>(ngll, phys, rect, a, s);
}

int force_blocks(int ngll, int nb) {
  const int epb = kWG / (ngll * ngll);
  return (nb + epb - 1) / epb;
}

}  // namespace insem_hip
