/*
 * insem_oracle.c -- CPU restatement of insem's explicit time-stepping hot
 * path.  TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); the
 * product never links or calls it.
 *
 * Every loop follows the reference's iteration order, summation order,
 * casts and expression forms, so that, compiled without FMA contraction, it
 * reproduces the reference's float build bit for bit:
 *   element kernels          cpp/src/cpu/kernels.cpp:19-431 (j outer, i inner, l ascending)
 *   point-wise physics       cpp/include/insem/physics/pointwise.hpp:12-57
 *   gather / assemble        cpp/include/insem/ops/dof_map.hpp:94-141
 *   time loop and boundary   cpp/src/solver/solver.cpp:67-87, 570-963
 *   treatment, sources,
 *   coupling, receivers,
 *   energy, power iteration
 * Each function below names the lines it restates.  OpenMP parallelises the
 * same loops the reference parallelises; results do not depend on the
 * thread count.  Built twice: liboracle.so computes in float (the
 * reference's float build), liboracle64.so in double (ORACLE_REAL=double,
 * the reference's double build) from the same float setup arrays.
 *
 * Pinned, independently of these formulas, by the tests in tests/:
 *   element kernels, gather/assemble  closed-form exactness, adjointness,
 *                                     mass = area (test_oracle_operator.py,
 *                                     test_setup.py)
 *   sources (force, moment tensor,    2D and 3D (axisymmetric m = 0, 1)
 *   monopole, dipole), receivers,     Green's functions
 *   particle velocity, m = 1 rotation (test_oracle_verification.py)
 *   Stacey (acoustic, elastic P/S,    normal-incidence echoes, monotone decay
 *   axisymmetric m = 0/1)             (test_oracle_verification.py)
 *   coupling                          normal-incidence R/T, fluid-solid
 *                                     reciprocity (Cartesian and m = 1)
 *   power iteration                   dense eigenvalues, the coupled stability
 *                                     limit (test_oracle_physics.py,
 *                                     test_oracle_verification.py)
 *   energy, time loop                 conservation to round-off, integrator
 *                                     identity (test_oracle_physics.py,
 *                                     test_oracle_fp64.py)
 * Parity unpinned: none of the functions; only the exact value of the
 * reference's stable_dt (its libstdc++ start vectors, generated by the host
 * library, not here) is fixed by bounds and the checks above alone.
 */
#include "insem_oracle.h"

/* Arithmetic and state type: float restates the reference's float build
 * (INSEM_REAL=float), double its double build (INSEM_REAL=double,
 * cpp/include/insem/core/config.hpp:19-23).  The setup arrays are float in
 * both (the double build reads float-rounded coefficients and computes in
 * double). */
#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL real;

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KSQRT_HALF 0.70710678118654752440

static __thread char g_err[512];

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char* oracle_last_error(void) { return g_err; }

int oracle_set_threads(int32_t n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

typedef struct {
  real *u, *v, *a, *f, *ul, *fl;
} dom_state;

typedef struct {
  double prev_ax, prev_az, vel_x, vel_z;
} pv_state;

struct oracle_ctx {
  const insem_hip_setup* s;
  dom_state* st;
  double dt;
  int n_steps, step, initialized, have_time;
  const real* amp; /* [n_sources][n_steps + 1], copied */
  real* amp_copy;
  double potential;
  /* traces: per receiver-quantity a growable [n_samples][ncomp] buffer */
  int n_traces;
  int* trace_ncomp;
  double** trace;
  int n_samples, cap_samples;
  pv_state* pv;
  double* energy;
  int n_energy, cap_energy;
};

static int dom_of(const oracle_ctx* c, int medium) {
  for (int i = 0; i < c->s->n_domains; ++i)
    if (c->s->domains[i].medium == medium) return i;
  return -1;
}

static size_t dom_len(const insem_hip_domain* d) { return (size_t)d->ncomp * (size_t)d->n_global; }

/* Float values in and out of the state arrays (a copy in the float build). */
static void from_float(real* dst, const float* src, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = (real)src[i];
}

static void to_float(float* dst, const real* src, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = (float)src[i];
}

/* ------------------------------------------------------- gather / assemble */

/* dof_map.hpp:94-120 */
static void gather(const insem_hip_domain* d, int npe, const real* g, real* l) {
  const size_t ng = (size_t)d->n_global;
  const int nc = d->ncomp;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < d->n_elements; ++e) {
    const int32_t* map = d->l2g + (size_t)e * npe;
    for (int c = 0; c < nc; ++c) {
      const real* gc = g + c * ng;
      real* lc = l + ((size_t)e * nc + c) * npe;
      for (int k = 0; k < npe; ++k) {
        const int32_t m = map[k];
        if (m >= 0) {
          lc[k] = gc[m];
        } else {
          const int32_t r = -1 - m;
          real s = (real)0;
          for (int32_t q = d->row_offsets[r]; q < d->row_offsets[r + 1]; ++q)
            s += (real)d->row_weights[q] * gc[d->row_dofs[q]];
          lc[k] = s;
        }
      }
    }
  }
}

/* dof_map.hpp:122-141 */
static void assemble(const insem_hip_domain* d, int npe, const real* l, real* g) {
  const size_t ng = (size_t)d->n_global;
  const int nc = d->ncomp;
  const int weighted = d->t_weights != NULL;
#pragma omp parallel for schedule(static)
  for (int gi = 0; gi < d->n_global; ++gi) {
    for (int c = 0; c < nc; ++c) {
      real sum = (real)0;
      for (int32_t s = d->t_offsets[gi]; s < d->t_offsets[gi + 1]; ++s) {
        const size_t slot = (size_t)d->t_slots[s];
        const size_t e = slot / (size_t)npe;
        const size_t k = slot - e * (size_t)npe;
        const real v = l[(e * nc + c) * npe + k];
        sum += weighted ? (real)d->t_weights[s] * v : v;
      }
      g[c * ng + gi] = sum;
    }
  }
}

/* ---------------------------------------------------------- element kernels */

/* kernels.cpp:40-51 and 168-179 (Dx = Dy on Cartesian blocks) */
static inline void ref_deriv(int ng, const float* Dx, const float* Dy, const real* u, int i, int j, real* u_xi,
                             real* u_eta) {
  real a = (real)0, b = (real)0;
  for (int l = 0; l < ng; ++l) {
    a += Dx[i * ng + l] * u[j * ng + l];
    b += Dy[j * ng + l] * u[l * ng + i];
  }
  *u_xi = a;
  *u_eta = b;
}

/* kernels.cpp:53-62 and 181-190 */
static inline real weak_div(int ng, const float* Dx, const float* Dy, const real* fx, const real* fe, int i, int j) {
  real s = (real)0;
  for (int l = 0; l < ng; ++l) {
    s += Dx[l * ng + i] * fx[j * ng + l];
    s += Dy[l * ng + j] * fe[l * ng + i];
  }
  return s;
}

/* pointwise.hpp:22-27 */
static inline void phys_grad(real a_xi, real a_eta, real xi_x, real xi_z, real eta_x, real eta_z, real* a_x,
                             real* a_z) {
  *a_x = a_xi * xi_x + a_eta * eta_x;
  *a_z = a_xi * xi_z + a_eta * eta_z;
}

/* pointwise.hpp:33-38 */
static inline void ref_flux(real wj, real qx, real qz, real xi_x, real xi_z, real eta_x, real eta_z, real* f_xi,
                            real* f_eta) {
  *f_xi = wj * (qx * xi_x + qz * xi_z);
  *f_eta = wj * (qx * eta_x + qz * eta_z);
}

/* kernels.cpp:18-38 */
static inline void load_metric(int kind, const float* metric, int b, int npe, int k, real* xi_x, real* xi_z,
                               real* eta_x, real* eta_z, real* jac) {
  if (kind == INSEM_KIND_RECT) {
    const float* m = metric + (size_t)b * 3;
    *xi_x = m[0];
    *xi_z = (real)0;
    *eta_x = (real)0;
    *eta_z = m[1];
    *jac = m[2];
  } else {
    const float* m = metric + (kind == INSEM_KIND_AFFINE ? (size_t)b * 5 : ((size_t)b * npe + k) * 5);
    *xi_x = m[0];
    *xi_z = m[1];
    *eta_x = m[2];
    *eta_z = m[3];
    *jac = m[4];
  }
}

#define MAXNG 9
#define MAXNPE 81

/* kernels.cpp:64-120 */
static void elastic_cart(const insem_hip_setup* S, const insem_hip_block* B, const real* u, real* f) {
  const int ng = S->ngll, npe = ng * ng, kind = B->kind;
  const float *D = S->D, *w2 = S->w2;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n_elements; ++b) {
    const size_t e = (size_t)B->elements[b];
    const real* ux = u + e * 2 * npe;
    const real* uz = ux + npe;
    const float* lambda = B->coef0 + (size_t)b * npe;
    const float* mu = B->coef1 + (size_t)b * npe;
    real f1x[MAXNPE], f1z[MAXNPE], f2x[MAXNPE], f2z[MAXNPE];
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        real ux_xi, ux_eta, uz_xi, uz_eta;
        ref_deriv(ng, D, D, ux, i, j, &ux_xi, &ux_eta);
        ref_deriv(ng, D, D, uz, i, j, &uz_xi, &uz_eta);
        real xi_x, xi_z, eta_x, eta_z, jac;
        load_metric(kind, B->metric, b, npe, k, &xi_x, &xi_z, &eta_x, &eta_z, &jac);
        real ux_x, ux_z, uz_x, uz_z;
        if (kind == INSEM_KIND_RECT) {
          ux_x = ux_xi * xi_x;
          ux_z = ux_eta * eta_z;
          uz_x = uz_xi * xi_x;
          uz_z = uz_eta * eta_z;
        } else {
          phys_grad(ux_xi, ux_eta, xi_x, xi_z, eta_x, eta_z, &ux_x, &ux_z);
          phys_grad(uz_xi, uz_eta, xi_x, xi_z, eta_x, eta_z, &uz_x, &uz_z);
        }
        /* pointwise.hpp:12-19 */
        const real div = ux_x + uz_z;
        const real sxx = lambda[k] * div + (real)2 * mu[k] * ux_x;
        const real szz = lambda[k] * div + (real)2 * mu[k] * uz_z;
        const real sxz = mu[k] * (ux_z + uz_x);
        const real wj = w2[k] * jac;
        if (kind == INSEM_KIND_RECT) {
          f1x[k] = wj * sxx * xi_x;
          f1z[k] = wj * sxz * xi_x;
          f2x[k] = wj * sxz * eta_z;
          f2z[k] = wj * szz * eta_z;
        } else {
          ref_flux(wj, sxx, sxz, xi_x, xi_z, eta_x, eta_z, &f1x[k], &f2x[k]);
          ref_flux(wj, sxz, szz, xi_x, xi_z, eta_x, eta_z, &f1z[k], &f2z[k]);
        }
      }
    real* fx = f + e * 2 * npe;
    real* fz = fx + npe;
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        fx[j * ng + i] = -weak_div(ng, D, D, f1x, f2x, i, j);
        fz[j * ng + i] = -weak_div(ng, D, D, f1z, f2z, i, j);
      }
  }
}

/* kernels.cpp:122-158 */
static void acoustic_cart(const insem_hip_setup* S, const insem_hip_block* B, const real* p, real* f) {
  const int ng = S->ngll, npe = ng * ng, kind = B->kind;
  const float *D = S->D, *w2 = S->w2;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n_elements; ++b) {
    const size_t e = (size_t)B->elements[b];
    const real* pe = p + e * npe;
    const float* inv_rho = B->coef0 + (size_t)b * npe;
    real f1[MAXNPE], f2[MAXNPE];
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        real p_xi, p_eta;
        ref_deriv(ng, D, D, pe, i, j, &p_xi, &p_eta);
        real xi_x, xi_z, eta_x, eta_z, jac;
        load_metric(kind, B->metric, b, npe, k, &xi_x, &xi_z, &eta_x, &eta_z, &jac);
        const real c = w2[k] * jac * inv_rho[k];
        if (kind == INSEM_KIND_RECT) {
          f1[k] = c * p_xi * xi_x * xi_x;
          f2[k] = c * p_eta * eta_z * eta_z;
        } else {
          real p_x, p_z;
          phys_grad(p_xi, p_eta, xi_x, xi_z, eta_x, eta_z, &p_x, &p_z);
          ref_flux(c, p_x, p_z, xi_x, xi_z, eta_x, eta_z, &f1[k], &f2[k]);
        }
      }
    real* fe = f + e * npe;
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) fe[j * ng + i] = -weak_div(ng, D, D, f1, f2, i, j);
  }
}

/* kernels.cpp:195-204 */
static void spread_axis(int ng, const float* Dx, const float* Dy, const float* m, const real* g, real* f) {
  for (int jq = 0; jq < ng; ++jq) {
    const int kq = jq * ng;
    const real xi_r = m[5 * kq], eta_r = m[5 * kq + 2];
    for (int ia = 0; ia < ng; ++ia) f[jq * ng + ia] -= g[jq] * xi_r * Dx[ia];
    for (int ja = 0; ja < ng; ++ja) f[ja * ng] -= g[jq] * eta_r * Dy[jq * ng + ja];
  }
}

/* kernels.cpp:206-263 */
static void elastic_axi0(const insem_hip_setup* S, const insem_hip_block* B, const real* u, real* f) {
  const int ng = S->ngll, npe = ng * ng;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n_elements; ++b) {
    const size_t e = (size_t)B->elements[b];
    const int axial = B->axial[b] != 0;
    const float* Dx = axial ? S->Dj : S->D;
    const float* Dy = S->D;
    const real* ur = u + e * 2 * npe;
    const real* uz = ur + npe;
    const size_t p0 = (size_t)b * npe;
    const float* m = B->metric + p0 * 5;
    real f1r[MAXNPE], f2r[MAXNPE], f1z[MAXNPE], f2z[MAXNPE], hoop_direct[MAXNPE], axis_g[MAXNG];
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        const size_t p = p0 + k;
        real ur_xi, ur_eta, uz_xi, uz_eta;
        ref_deriv(ng, Dx, Dy, ur, i, j, &ur_xi, &ur_eta);
        ref_deriv(ng, Dx, Dy, uz, i, j, &uz_xi, &uz_eta);
        const real xi_r = m[5 * k], xi_z = m[5 * k + 1], eta_r = m[5 * k + 2], eta_z = m[5 * k + 3];
        real ur_r, ur_z, uz_r, uz_z;
        phys_grad(ur_xi, ur_eta, xi_r, xi_z, eta_r, eta_z, &ur_r, &ur_z);
        phys_grad(uz_xi, uz_eta, xi_r, xi_z, eta_r, eta_z, &uz_r, &uz_z);
        const real r = B->r[p];
        const int on_axis = r == (real)0;
        const real e_tt = on_axis ? ur_r : ur[k] / r;
        const real lambda = B->coef0[p], mu = B->coef1[p];
        const real tr = ur_r + uz_z + e_tt;
        const real srr = lambda * tr + 2 * mu * ur_r;
        const real szz = lambda * tr + 2 * mu * uz_z;
        const real stt = lambda * tr + 2 * mu * e_tt;
        const real srz = mu * (ur_z + uz_r);
        const real w = B->wr[p];
        ref_flux(w, srr, srz, xi_r, xi_z, eta_r, eta_z, &f1r[k], &f2r[k]);
        ref_flux(w, srz, szz, xi_r, xi_z, eta_r, eta_z, &f1z[k], &f2z[k]);
        if (axial && i == 0) {
          hoop_direct[k] = (real)0;
          axis_g[j] = w * stt;
        } else {
          hoop_direct[k] = B->hoop[p] * stt;
        }
      }
    real* fr = f + e * 2 * npe;
    real* fz = fr + npe;
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        fr[k] = -(weak_div(ng, Dx, Dy, f1r, f2r, i, j) + hoop_direct[k]);
        fz[k] = -weak_div(ng, Dx, Dy, f1z, f2z, i, j);
      }
    if (axial) spread_axis(ng, Dx, Dy, m, axis_g, fr);
  }
}

/* kernels.cpp:265-294 */
static void acoustic_axi0(const insem_hip_setup* S, const insem_hip_block* B, const real* pin, real* f) {
  const int ng = S->ngll, npe = ng * ng;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n_elements; ++b) {
    const size_t e = (size_t)B->elements[b];
    const float* Dx = B->axial[b] ? S->Dj : S->D;
    const float* Dy = S->D;
    const real* pe = pin + e * npe;
    const size_t p0 = (size_t)b * npe;
    const float* m = B->metric + p0 * 5;
    real f1[MAXNPE], f2[MAXNPE];
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        real p_xi, p_eta;
        ref_deriv(ng, Dx, Dy, pe, i, j, &p_xi, &p_eta);
        const real xi_r = m[5 * k], xi_z = m[5 * k + 1], eta_r = m[5 * k + 2], eta_z = m[5 * k + 3];
        real p_r, p_z;
        phys_grad(p_xi, p_eta, xi_r, xi_z, eta_r, eta_z, &p_r, &p_z);
        ref_flux(B->wr[p0 + k] * B->coef0[p0 + k], p_r, p_z, xi_r, xi_z, eta_r, eta_z, &f1[k], &f2[k]);
      }
    real* fe = f + e * npe;
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) fe[j * ng + i] = -weak_div(ng, Dx, Dy, f1, f2, i, j);
  }
}

/* kernels.cpp:305-385 */
static void elastic_axi1(const insem_hip_setup* S, const insem_hip_block* B, const real* u, real* f) {
  const int ng = S->ngll, npe = ng * ng;
  const real s2 = (real)1.41421356237309504880;
  const real is2 = (real)0.70710678118654752440;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n_elements; ++b) {
    const size_t e = (size_t)B->elements[b];
    const int axial = B->axial[b] != 0;
    const float* Dx = axial ? S->Dj : S->D;
    const float* Dy = S->D;
    const real* up = u + e * 3 * npe;
    const real* um = up + npe;
    const real* uz = um + npe;
    const size_t p0 = (size_t)b * npe;
    const float* m = B->metric + p0 * 5;
    real f1p[MAXNPE], f2p[MAXNPE], f1m[MAXNPE], f2m[MAXNPE], f1z[MAXNPE], f2z[MAXNPE];
    real direct_p[MAXNPE], direct_z[MAXNPE], axis_p[MAXNG], axis_z[MAXNG];
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        const size_t p = p0 + k;
        real up_xi, up_eta, um_xi, um_eta, uz_xi, uz_eta;
        ref_deriv(ng, Dx, Dy, up, i, j, &up_xi, &up_eta);
        ref_deriv(ng, Dx, Dy, um, i, j, &um_xi, &um_eta);
        ref_deriv(ng, Dx, Dy, uz, i, j, &uz_xi, &uz_eta);
        const real xi_r = m[5 * k], xi_z = m[5 * k + 1], eta_r = m[5 * k + 2], eta_z = m[5 * k + 3];
        real up_r, up_z, um_r, um_z, uz_r, uz_z;
        phys_grad(up_xi, up_eta, xi_r, xi_z, eta_r, eta_z, &up_r, &up_z);
        phys_grad(um_xi, um_eta, xi_r, xi_z, eta_r, eta_z, &um_r, &um_z);
        phys_grad(uz_xi, uz_eta, xi_r, xi_z, eta_r, eta_z, &uz_r, &uz_z);
        const real r = B->r[p];
        const int on_axis = r == (real)0;
        const real up_over_r = on_axis ? up_r : up[k] / r;
        const real uz_over_r = on_axis ? uz_r : uz[k] / r;
        const real ur_r = (up_r + um_r) * is2, ur_z = (up_z + um_z) * is2;
        const real ut_r = (up_r - um_r) * is2, ut_z = (up_z - um_z) * is2;
        const real e_tt = s2 * up_over_r;
        const real lambda = B->coef0[p], mu = B->coef1[p];
        const real tr = ur_r + uz_z + e_tt;
        const real srr = lambda * tr + 2 * mu * ur_r;
        const real szz = lambda * tr + 2 * mu * uz_z;
        const real stt = lambda * tr + 2 * mu * e_tt;
        const real srz = mu * (ur_z + uz_r);
        const real srt = mu * (ut_r - s2 * up_over_r);
        const real stz = mu * (ut_z - uz_over_r);
        const real w = B->wr[p];
        ref_flux(w, (srr + srt) * is2, (srz + stz) * is2, xi_r, xi_z, eta_r, eta_z, &f1p[k], &f2p[k]);
        ref_flux(w, (srr - srt) * is2, (srz - stz) * is2, xi_r, xi_z, eta_r, eta_z, &f1m[k], &f2m[k]);
        ref_flux(w, srz, szz, xi_r, xi_z, eta_r, eta_z, &f1z[k], &f2z[k]);
        const real cp = s2 * (stt - srt), cz = -stz;
        if (axial && i == 0) {
          direct_p[k] = direct_z[k] = (real)0;
          axis_p[j] = w * cp;
          axis_z[j] = w * cz;
        } else {
          direct_p[k] = B->hoop[p] * cp;
          direct_z[k] = B->hoop[p] * cz;
        }
      }
    real* fp = f + e * 3 * npe;
    real* fm = fp + npe;
    real* fz = fm + npe;
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        fp[k] = -(weak_div(ng, Dx, Dy, f1p, f2p, i, j) + direct_p[k]);
        fm[k] = -weak_div(ng, Dx, Dy, f1m, f2m, i, j);
        fz[k] = -(weak_div(ng, Dx, Dy, f1z, f2z, i, j) + direct_z[k]);
      }
    if (axial) {
      spread_axis(ng, Dx, Dy, m, axis_p, fp);
      spread_axis(ng, Dx, Dy, m, axis_z, fz);
    }
  }
}

/* kernels.cpp:387-431 */
static void acoustic_axi1(const insem_hip_setup* S, const insem_hip_block* B, const real* pin, real* f) {
  const int ng = S->ngll, npe = ng * ng;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n_elements; ++b) {
    const size_t e = (size_t)B->elements[b];
    const int axial = B->axial[b] != 0;
    const float* Dx = axial ? S->Dj : S->D;
    const float* Dy = S->D;
    const real* pe = pin + e * npe;
    const size_t p0 = (size_t)b * npe;
    const float* m = B->metric + p0 * 5;
    real f1[MAXNPE], f2[MAXNPE], direct[MAXNPE], axis_g[MAXNG];
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        const size_t p = p0 + k;
        real p_xi, p_eta;
        ref_deriv(ng, Dx, Dy, pe, i, j, &p_xi, &p_eta);
        const real xi_r = m[5 * k], xi_z = m[5 * k + 1], eta_r = m[5 * k + 2], eta_z = m[5 * k + 3];
        real p_r, p_z;
        phys_grad(p_xi, p_eta, xi_r, xi_z, eta_r, eta_z, &p_r, &p_z);
        const real inv_rho = B->coef0[p];
        ref_flux(B->wr[p] * inv_rho, p_r, p_z, xi_r, xi_z, eta_r, eta_z, &f1[k], &f2[k]);
        if (axial && i == 0) {
          direct[k] = (real)0;
          axis_g[j] = B->wr[p] * inv_rho * p_r;
        } else {
          direct[k] = B->hoop[p] * inv_rho * pe[k] / B->r[p];
        }
      }
    real* fe = f + e * npe;
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        const int k = j * ng + i;
        fe[k] = -(weak_div(ng, Dx, Dy, f1, f2, i, j) + direct[k]);
      }
    if (axial) spread_axis(ng, Dx, Dy, m, axis_g, fe);
  }
}

/* kernels.cpp:435-495: dispatch by physics, kind and azimuthal order */
static int run_block(const insem_hip_setup* S, const insem_hip_domain* d, const insem_hip_block* B, const real* ul,
                     real* fl) {
  const int elastic = d->medium == INSEM_MEDIUM_ELASTIC;
  if (S->ngll < 2 || S->ngll > 9) return fail(INSEM_HIP_EINVAL, "no kernel for %d GLL points (supported: 2..9)", S->ngll);
  if (!(B->azimuthal_order == 0 || (B->azimuthal_order == 1 && B->kind == INSEM_KIND_AXISYM)))
    return fail(INSEM_HIP_EINVAL, "no kernel for azimuthal order %d on kind %d elements", B->azimuthal_order, B->kind);
  if (B->kind == INSEM_KIND_AXISYM) {
    if (B->azimuthal_order == 1) {
      if (elastic)
        elastic_axi1(S, B, ul, fl);
      else
        acoustic_axi1(S, B, ul, fl);
    } else if (elastic) {
      elastic_axi0(S, B, ul, fl);
    } else {
      acoustic_axi0(S, B, ul, fl);
    }
  } else if (elastic) {
    elastic_cart(S, B, ul, fl);
  } else {
    acoustic_cart(S, B, ul, fl);
  }
  return INSEM_HIP_OK;
}

/* solver.cpp:82-87 */
static double dot_fd(const real* a, const real* b, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
  return s;
}

/* solver.cpp:570-581 */
static double internal_forces(oracle_ctx* c, int di, const real* u, int energy) {
  const insem_hip_setup* S = c->s;
  const insem_hip_domain* d = &S->domains[di];
  dom_state* st = &c->st[di];
  const int npe = S->ngll * S->ngll;
  gather(d, npe, u, st->ul);
  for (int b = 0; b < d->n_blocks; ++b) run_block(S, d, &d->blocks[b], st->ul, st->fl);
  assemble(d, npe, st->fl, st->f);
  return energy ? dot_fd(u, st->f, dom_len(d)) : 0.0;
}

/* solver.hpp:57-65 */
static void apply_dirichlet(const insem_hip_domain* d, real* x) {
  const size_t ng = (size_t)d->n_global;
  for (int c = 0; c < d->ncomp; ++c)
    for (int i = 0; i < d->n_dirichlet; ++i) x[c * ng + d->dirichlet_points[i]] = (real)0;
  for (int ci = 0; ci < d->n_axis_components; ++ci) {
    const int c = d->axis_components[ci];
    for (int i = 0; i < d->n_axis; ++i) x[c * ng + d->axis_points[i]] = (real)0;
  }
}

/* solver.cpp:693-727 */
static void add_coupling(oracle_ctx* c, int di) {
  const insem_hip_setup* S = c->s;
  const insem_hip_coupling* cp = &S->coupling;
  const int n = cp->n_points;
  if (n == 0) return;
  const insem_hip_domain* d = &S->domains[di];
  const int pm = d->azimuthal_order == 1;
  const real h = (real)KSQRT_HALF;
  if (d->medium == INSEM_MEDIUM_ELASTIC) {
    const int fi = dom_of(c, INSEM_MEDIUM_ACOUSTIC);
    const real* pu = c->st[fi].u;
    real* f = c->st[di].f;
    const size_t ng = (size_t)d->n_global;
    const size_t zc = pm ? 2 * ng : ng;
    for (int i = 0; i < n; ++i) {
      const real p = pu[cp->fluid_points[i]];
      const size_t g = (size_t)cp->solid_points[i];
      if (pm) {
        f[g] += h * cp->cx[i] * p;
        f[ng + g] += h * cp->cx[i] * p;
      } else {
        f[g] += cp->cx[i] * p;
      }
      f[zc + g] += cp->cz[i] * p;
    }
  } else {
    const int si = dom_of(c, INSEM_MEDIUM_ELASTIC);
    const real* a = c->st[si].a;
    const size_t ngs = (size_t)S->domains[si].n_global;
    const size_t zc = pm ? 2 * ngs : ngs;
    real* f = c->st[di].f;
    for (int i = 0; i < n; ++i) {
      const size_t gs = (size_t)cp->solid_points[i];
      const real ar = pm ? h * (a[gs] + a[ngs + gs]) : a[gs];
      f[cp->fluid_points[i]] -= cp->cx[i] * ar + cp->cz[i] * a[zc + gs];
    }
  }
}

/* solver.cpp:67-80 */
static void solve_spd(int n, double* A, double* b) {
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) {
      const double l = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
      b[i] -= l * b[k];
    }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
    b[k] = s / A[k * n + k];
  }
}

/* solver.cpp:729-762 */
static void solve_acceleration(oracle_ctx* c, int di, int damping_implicit) {
  const insem_hip_domain* d = &c->s->domains[di];
  dom_state* st = &c->st[di];
  const size_t ng = (size_t)d->n_global;
  const int nc = d->ncomp;
  const float* im = d->inv_mass;
#pragma omp parallel for schedule(static)
  for (long g = 0; g < (long)ng; ++g)
    for (int cc = 0; cc < nc; ++cc) st->a[cc * ng + g] = st->f[cc * ng + g] * im[g];
  const double h = damping_implicit ? 0.5 * c->dt : 0.0;
  for (int i = 0; i < d->n_absorbing; ++i) {
    const size_t g = (size_t)d->absorbing_points[i];
    const float* cf = d->absorbing_coef + (size_t)nc * nc * i;
    double A[9], rhs[3];
    for (int r = 0; r < nc; ++r) {
      rhs[r] = st->f[r * ng + g];
      for (int s = 0; s < nc; ++s) {
        rhs[r] -= (double)cf[r * nc + s] * st->v[s * ng + g];
        A[r * nc + s] = h * cf[r * nc + s] + (r == s ? (double)d->mass[g] : 0.0);
      }
    }
    solve_spd(nc, A, rhs);
    for (int r = 0; r < nc; ++r) st->a[r * ng + g] = (real)rhs[r];
  }
  apply_dirichlet(d, st->a);
}

/* solver.cpp:669-691 */
static void compute_acceleration(oracle_ctx* c, int energy, int damping_implicit) {
  const insem_hip_setup* S = c->s;
  c->potential = 0.0;
  const int order[2] = {INSEM_MEDIUM_ELASTIC, INSEM_MEDIUM_ACOUSTIC};
  for (int o = 0; o < 2; ++o) {
    const int di = dom_of(c, order[o]);
    if (di < 0) continue;
    const insem_hip_domain* d = &S->domains[di];
    dom_state* st = &c->st[di];
    c->potential -= 0.5 * internal_forces(c, di, st->u, energy);
    for (int s = 0; s < S->n_sources; ++s) {
      const insem_hip_source* src = &S->sources[s];
      if (src->domain != di) continue;
      const real amp = c->amp[(size_t)s * (c->n_steps + 1) + c->step];
      const size_t ng = (size_t)d->n_global, np = (size_t)src->n_points;
      for (int cc = 0; cc < d->ncomp; ++cc)
        for (size_t i = 0; i < np; ++i) st->f[cc * ng + src->points[i]] += src->coef[cc * np + i] * amp;
    }
    add_coupling(c, di);
    solve_acceleration(c, di, damping_implicit);
  }
}

/* solver.cpp:900-915 */
static double kinetic_energy(const oracle_ctx* c) {
  const double cc = c->dt * c->dt / 8.0;
  double kin = 0.0;
  for (int di = 0; di < c->s->n_domains; ++di) {
    const insem_hip_domain* d = &c->s->domains[di];
    const dom_state* st = &c->st[di];
    const size_t ng = (size_t)d->n_global;
    for (int comp = 0; comp < d->ncomp; ++comp)
      for (size_t g = 0; g < ng; ++g) {
        const double v = st->v[comp * ng + g], a = st->a[comp * ng + g];
        kin += d->mass[g] * (0.5 * v * v - cc * a * a);
      }
  }
  return kin;
}

static int grow_samples(oracle_ctx* c) {
  if (c->n_samples < c->cap_samples) return 0;
  const int cap = c->cap_samples ? 2 * c->cap_samples : 64;
  for (int t = 0; t < c->n_traces; ++t) {
    double* p = (double*)realloc(c->trace[t], sizeof(double) * (size_t)cap * c->trace_ncomp[t]);
    if (!p) return -1;
    c->trace[t] = p;
  }
  c->cap_samples = cap;
  return 0;
}

static double interp(const insem_hip_receiver* r, const real* field, size_t ng, int comp) {
  double s = 0.0;
  for (int i = 0; i < r->n_points; ++i) s += (double)r->value[i] * field[comp * ng + r->points[i]];
  return s;
}

/* solver.cpp:780-811 */
static void record(oracle_ctx* c) {
  const insem_hip_setup* S = c->s;
  if (grow_samples(c)) return;
  const int n = c->n_samples;
  int t = 0;
  for (int ri = 0; ri < S->n_receivers; ++ri) {
    const insem_hip_receiver* r = &S->receivers[ri];
    const insem_hip_domain* d = &S->domains[r->domain];
    const dom_state* st = &c->st[r->domain];
    const size_t ng = (size_t)d->n_global;
    for (int qi = 0; qi < r->n_quantities; ++qi, ++t) {
      const int q = r->quantities[qi];
      double* out = c->trace[t] + (size_t)n * c->trace_ncomp[t];
      if (q == INSEM_Q_PARTICLE_VELOCITY) {
        out[0] = c->pv[ri].vel_x;
        out[1] = c->pv[ri].vel_z;
        continue;
      }
      const real* field = q == INSEM_Q_VELOCITY ? st->v : q == INSEM_Q_ACCELERATION ? st->a : st->u;
      if (d->medium == INSEM_MEDIUM_ELASTIC && d->azimuthal_order == 1) {
        const double up = interp(r, field, ng, 0), um = interp(r, field, ng, 1);
        out[0] = KSQRT_HALF * (up + um);
        out[1] = KSQRT_HALF * (up - um);
        out[2] = interp(r, field, ng, 2);
        continue;
      }
      for (int cc = 0; cc < c->trace_ncomp[t]; ++cc) out[cc] = interp(r, field, ng, cc);
    }
  }
  c->n_samples = n + 1;
}

static void push_energy(oracle_ctx* c, double t, double kin, double pot) {
  if (c->n_energy == c->cap_energy) {
    const int cap = c->cap_energy ? 2 * c->cap_energy : 64;
    double* p = (double*)realloc(c->energy, sizeof(double) * 3 * (size_t)cap);
    if (!p) return;
    c->energy = p;
    c->cap_energy = cap;
  }
  c->energy[3 * c->n_energy] = t;
  c->energy[3 * c->n_energy + 1] = kin;
  c->energy[3 * c->n_energy + 2] = pot;
  ++c->n_energy;
}

static void pressure_gradient(const oracle_ctx* c, int ri, double* gx, double* gz) {
  const insem_hip_receiver* r = &c->s->receivers[ri];
  const real* u = c->st[r->domain].u;
  double x = 0.0, z = 0.0;
  for (int i = 0; i < r->n_points; ++i) {
    x += (double)r->d_dx[i] * u[r->points[i]];
    z += (double)r->d_dz[i] * u[r->points[i]];
  }
  *gx = x;
  *gz = z;
}

/* solver.cpp:878-898 */
static void initialize(oracle_ctx* c) {
  const int energy = c->s->energy_every > 0;
  compute_acceleration(c, energy, 0);
  for (int ri = 0; ri < c->s->n_receivers; ++ri) {
    const insem_hip_receiver* r = &c->s->receivers[ri];
    if (c->s->domains[r->domain].medium != INSEM_MEDIUM_ACOUSTIC) continue;
    double gx, gz;
    pressure_gradient(c, ri, &gx, &gz);
    c->pv[ri].prev_ax = -gx / r->rho;
    c->pv[ri].prev_az = -gz / r->rho;
  }
  c->initialized = 1;
  record(c);
  if (energy) push_energy(c, 0.0, kinetic_energy(c), c->potential);
}

/* ---------------------------------------------------------------- public */

int oracle_create(const insem_hip_setup* s, oracle_ctx** out) {
  if (!s || !out) return fail(INSEM_HIP_EINVAL, "null argument");
  if (s->ngll < 2 || s->ngll > 9) return fail(INSEM_HIP_EINVAL, "no kernel for %d GLL points (supported: 2..9)", s->ngll);
  oracle_ctx* c = (oracle_ctx*)calloc(1, sizeof *c);
  c->s = s;
  const size_t npe = (size_t)s->ngll * s->ngll;
  c->st = (dom_state*)calloc((size_t)s->n_domains, sizeof(dom_state));
  for (int di = 0; di < s->n_domains; ++di) {
    const insem_hip_domain* d = &s->domains[di];
    const size_t n = dom_len(d), nl = (size_t)d->ncomp * d->n_elements * npe;
    c->st[di].u = (real*)calloc(n, sizeof(real));
    c->st[di].v = (real*)calloc(n, sizeof(real));
    c->st[di].a = (real*)calloc(n, sizeof(real));
    c->st[di].f = (real*)calloc(n, sizeof(real));
    c->st[di].ul = (real*)calloc(nl, sizeof(real));
    c->st[di].fl = (real*)calloc(nl, sizeof(real));
  }
  int nt = 0;
  for (int r = 0; r < s->n_receivers; ++r) nt += s->receivers[r].n_quantities;
  c->n_traces = nt;
  c->trace_ncomp = (int*)calloc((size_t)nt + 1, sizeof(int));
  c->trace = (double**)calloc((size_t)nt + 1, sizeof(double*));
  int t = 0;
  for (int r = 0; r < s->n_receivers; ++r) {
    const insem_hip_receiver* rc = &s->receivers[r];
    const insem_hip_domain* d = &s->domains[rc->domain];
    for (int q = 0; q < rc->n_quantities; ++q, ++t) {
      if (rc->quantities[q] == INSEM_Q_PARTICLE_VELOCITY)
        c->trace_ncomp[t] = 2;
      else
        c->trace_ncomp[t] = d->ncomp;
    }
  }
  c->pv = (pv_state*)calloc((size_t)s->n_receivers + 1, sizeof(pv_state));
  *out = c;
  return INSEM_HIP_OK;
}

int oracle_destroy(oracle_ctx* c) {
  if (!c) return INSEM_HIP_OK;
  for (int di = 0; di < c->s->n_domains; ++di) {
    free(c->st[di].u);
    free(c->st[di].v);
    free(c->st[di].a);
    free(c->st[di].f);
    free(c->st[di].ul);
    free(c->st[di].fl);
  }
  free(c->st);
  for (int t = 0; t < c->n_traces; ++t) free(c->trace[t]);
  free(c->trace);
  free(c->trace_ncomp);
  free(c->pv);
  free(c->energy);
  free(c->amp_copy);
  free(c);
  return INSEM_HIP_OK;
}

int oracle_set_time(oracle_ctx* c, double dt, int32_t n_steps, const float* amplitudes) {
  if (!c || dt <= 0.0 || n_steps < 0) return fail(INSEM_HIP_EINVAL, "invalid time step or step count");
  if (c->initialized) return fail(INSEM_HIP_ESTATE, "the time step must be set before the first step");
  c->dt = dt;
  c->n_steps = n_steps;
  free(c->amp_copy);
  const size_t n = (size_t)c->s->n_sources * ((size_t)n_steps + 1);
  c->amp_copy = (real*)calloc(n + 1, sizeof(real));
  if (n && amplitudes) from_float(c->amp_copy, amplitudes, n);
  c->amp = c->amp_copy;
  c->have_time = 1;
  return INSEM_HIP_OK;
}

/* solver.cpp:583-667 */
int oracle_lambda_max(oracle_ctx* c, const float* start, double* lambda_out) {
  if (!c || !start || !lambda_out) return fail(INSEM_HIP_EINVAL, "null argument");
  const insem_hip_setup* S = c->s;
  double lambda = 0.0;
  size_t off = 0;
  for (int di = 0; di < S->n_domains; ++di) {
    const insem_hip_domain* d = &S->domains[di];
    dom_state* st = &c->st[di];
    const size_t ng = (size_t)d->n_global, len = dom_len(d);
    real* x = (real*)malloc(len * sizeof(real));
    from_float(x, start + off, len);
    off += len;
    for (int it = 0; it < 100; ++it) {
      apply_dirichlet(d, x);
      const double xkx = -internal_forces(c, di, x, 1);
      double xmx = 0.0, norm = 0.0;
      for (int cc = 0; cc < d->ncomp; ++cc)
        for (size_t g = 0; g < ng; ++g) xmx += (double)d->mass[g] * x[cc * ng + g] * x[cc * ng + g];
      if (xmx > 0.0 && xkx / xmx > lambda) lambda = xkx / xmx;
      for (int cc = 0; cc < d->ncomp; ++cc)
        for (size_t g = 0; g < ng; ++g) {
          const double y = -(double)st->f[cc * ng + g] * d->inv_mass[g];
          x[cc * ng + g] = (real)y;
          norm += y * y;
        }
      if (norm == 0.0) break;
      const double scale = 1.0 / sqrt(norm);
      for (size_t i = 0; i < len; ++i) x[i] = (real)(x[i] * scale);
    }
    memset(st->f, 0, len * sizeof(real));
    free(x);
  }
  if (S->coupling.n_points > 0) {
    const int si = dom_of(c, INSEM_MEDIUM_ELASTIC), fi = dom_of(c, INSEM_MEDIUM_ACOUSTIC);
    const int order[2] = {si, fi};
    real* saved[4];
    const size_t len_s = dom_len(&S->domains[si]), len_f = dom_len(&S->domains[fi]);
    saved[0] = (real*)malloc(len_s * sizeof(real));
    saved[1] = (real*)malloc(len_f * sizeof(real));
    saved[2] = (real*)malloc(len_s * sizeof(real));
    saved[3] = (real*)malloc(len_f * sizeof(real));
    memcpy(saved[0], c->st[si].u, len_s * sizeof(real));
    memcpy(saved[1], c->st[fi].u, len_f * sizeof(real));
    memcpy(saved[2], c->st[si].a, len_s * sizeof(real));
    memcpy(saved[3], c->st[fi].a, len_f * sizeof(real));
    from_float(c->st[si].u, start + off, len_s);
    off += len_s;
    from_float(c->st[fi].u, start + off, len_f);
    double lam = 0.0;
    for (int it = 0; it < 120; ++it) {
      for (int o = 0; o < 2; ++o) apply_dirichlet(&S->domains[order[o]], c->st[order[o]].u);
      double nx = 0.0, ny = 0.0;
      for (int o = 0; o < 2; ++o) {
        const insem_hip_domain* d = &S->domains[order[o]];
        const size_t ng = (size_t)d->n_global;
        double s = 0.0;
        for (int cc = 0; cc < d->ncomp; ++cc)
          for (size_t g = 0; g < ng; ++g)
            s += (double)d->mass[g] * c->st[order[o]].u[cc * ng + g] * c->st[order[o]].u[cc * ng + g];
        nx += s;
      }
      for (int o = 0; o < 2; ++o) {
        const int di = order[o];
        const insem_hip_domain* d = &S->domains[di];
        dom_state* st = &c->st[di];
        internal_forces(c, di, st->u, 0);
        add_coupling(c, di);
        const size_t ng = (size_t)d->n_global;
        for (int cc = 0; cc < d->ncomp; ++cc)
          for (size_t g = 0; g < ng; ++g) st->a[cc * ng + g] = st->f[cc * ng + g] * d->inv_mass[g];
        apply_dirichlet(d, st->a);
      }
      for (int o = 0; o < 2; ++o) {
        const insem_hip_domain* d = &S->domains[order[o]];
        const size_t ng = (size_t)d->n_global;
        double s = 0.0;
        for (int cc = 0; cc < d->ncomp; ++cc)
          for (size_t g = 0; g < ng; ++g)
            s += (double)d->mass[g] * c->st[order[o]].a[cc * ng + g] * c->st[order[o]].a[cc * ng + g];
        ny += s;
      }
      if (nx <= 0.0 || ny <= 0.0) break;
      if (it >= 60 && sqrt(ny / nx) > lam) lam = sqrt(ny / nx);
      const double scale = -1.0 / sqrt(ny);
      for (int o = 0; o < 2; ++o) {
        dom_state* st = &c->st[order[o]];
        const size_t len = dom_len(&S->domains[order[o]]);
        for (size_t i = 0; i < len; ++i) st->u[i] = (real)(st->a[i] * scale);
      }
    }
    memcpy(c->st[si].u, saved[0], len_s * sizeof(real));
    memcpy(c->st[fi].u, saved[1], len_f * sizeof(real));
    memcpy(c->st[si].a, saved[2], len_s * sizeof(real));
    memcpy(c->st[fi].a, saved[3], len_f * sizeof(real));
    memset(c->st[si].f, 0, len_s * sizeof(real));
    memset(c->st[fi].f, 0, len_f * sizeof(real));
    for (int k = 0; k < 4; ++k) free(saved[k]);
    if (lam > lambda) lambda = lam;
  }
  *lambda_out = lambda;
  return INSEM_HIP_OK;
}

/* solver.cpp:764-778 */
int oracle_set_state(oracle_ctx* c, int32_t di, const float* u, const float* v) {
  if (!c || di < 0 || di >= c->s->n_domains) return fail(INSEM_HIP_EINVAL, "invalid domain");
  if (c->initialized) return fail(INSEM_HIP_ESTATE, "initial conditions must be set before the first time step");
  const insem_hip_domain* d = &c->s->domains[di];
  const size_t n = dom_len(d);
  from_float(c->st[di].u, u, n);
  from_float(c->st[di].v, v, n);
  apply_dirichlet(d, c->st[di].u);
  apply_dirichlet(d, c->st[di].v);
  return INSEM_HIP_OK;
}

int oracle_get_field(oracle_ctx* c, int32_t di, int32_t field, float* out) {
  if (!c || di < 0 || di >= c->s->n_domains) return fail(INSEM_HIP_EINVAL, "invalid domain");
  const dom_state* st = &c->st[di];
  const real* src = field == INSEM_FIELD_U ? st->u : field == INSEM_FIELD_V ? st->v : field == INSEM_FIELD_A ? st->a : st->f;
  to_float(out, src, dom_len(&c->s->domains[di]));
  return INSEM_HIP_OK;
}

/* solver.cpp:917-963 */
int oracle_advance(oracle_ctx* c, int32_t n) {
  if (!c) return fail(INSEM_HIP_EINVAL, "null context");
  if (!c->have_time) return fail(INSEM_HIP_ESTATE, "set_time has not been called");
  if (!c->initialized) initialize(c);
  const insem_hip_setup* S = c->s;
  const double dt = c->dt;
  const real rdt = (real)dt;
  const real half_dt = (real)(0.5 * dt);
  const real half_dt2 = (real)(0.5 * dt * dt);
  for (int s = 0; s < n && c->step < c->n_steps; ++s) {
    for (int di = 0; di < S->n_domains; ++di) {
      real* u = c->st[di].u;
      real* v = c->st[di].v;
      const real* a = c->st[di].a;
      const long len = (long)dom_len(&S->domains[di]);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < len; ++i) {
        u[i] += rdt * v[i] + half_dt2 * a[i];
        v[i] += half_dt * a[i];
      }
    }
    ++c->step;
    const int energy = S->energy_every > 0 && c->step % S->energy_every == 0;
    compute_acceleration(c, energy, 1);
    for (int di = 0; di < S->n_domains; ++di) {
      real* v = c->st[di].v;
      const real* a = c->st[di].a;
      const long len = (long)dom_len(&S->domains[di]);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < len; ++i) v[i] += half_dt * a[i];
    }
    for (int ri = 0; ri < S->n_receivers; ++ri) {
      const insem_hip_receiver* r = &S->receivers[ri];
      if (S->domains[r->domain].medium != INSEM_MEDIUM_ACOUSTIC) continue;
      double gx, gz;
      pressure_gradient(c, ri, &gx, &gz);
      const double ax = -gx / r->rho, az = -gz / r->rho;
      c->pv[ri].vel_x += 0.5 * dt * (c->pv[ri].prev_ax + ax);
      c->pv[ri].vel_z += 0.5 * dt * (c->pv[ri].prev_az + az);
      c->pv[ri].prev_ax = ax;
      c->pv[ri].prev_az = az;
    }
    if (c->step % S->record_every == 0) record(c);
    if (energy) push_energy(c, c->step * dt, kinetic_energy(c), c->potential);
  }
  return INSEM_HIP_OK;
}

int oracle_step(oracle_ctx* c, int32_t* step) {
  if (!c || !step) return fail(INSEM_HIP_EINVAL, "null argument");
  *step = c->step;
  return INSEM_HIP_OK;
}

int oracle_n_samples(oracle_ctx* c, int32_t* n) {
  if (!c || !n) return fail(INSEM_HIP_EINVAL, "null argument");
  *n = c->n_samples;
  return INSEM_HIP_OK;
}

int oracle_read_traces(oracle_ctx* c, double* out) {
  if (!c || !out) return fail(INSEM_HIP_EINVAL, "null argument");
  size_t o = 0;
  for (int t = 0; t < c->n_traces; ++t) {
    const size_t n = (size_t)c->n_samples * c->trace_ncomp[t];
    if (n) memcpy(out + o, c->trace[t], n * sizeof(double));
    o += n;
  }
  return INSEM_HIP_OK;
}

int oracle_n_energy(oracle_ctx* c, int32_t* n) {
  if (!c || !n) return fail(INSEM_HIP_EINVAL, "null argument");
  *n = c->n_energy;
  return INSEM_HIP_OK;
}

int oracle_read_energy(oracle_ctx* c, double* out) {
  if (!c || !out) return fail(INSEM_HIP_EINVAL, "null argument");
  if (c->n_energy) memcpy(out, c->energy, sizeof(double) * 3 * (size_t)c->n_energy);
  return INSEM_HIP_OK;
}

int oracle_internal_forces(oracle_ctx* c, int32_t di, const float* u, float* f, double* dot) {
  if (!c || di < 0 || di >= c->s->n_domains) return fail(INSEM_HIP_EINVAL, "invalid domain");
  const size_t n = dom_len(&c->s->domains[di]);
  real* x = (real*)calloc(n + 1, sizeof(real));
  from_float(x, u, n);
  const double d = internal_forces(c, di, x, dot != NULL);
  free(x);
  to_float(f, c->st[di].f, n);
  if (dot) *dot = d;
  return INSEM_HIP_OK;
}

int oracle_block_forces(oracle_ctx* c, int32_t di, int32_t bi, const float* ul, float* fl) {
  if (!c || di < 0 || di >= c->s->n_domains) return fail(INSEM_HIP_EINVAL, "invalid domain");
  const insem_hip_domain* d = &c->s->domains[di];
  if (bi < 0 || bi >= d->n_blocks) return fail(INSEM_HIP_EINVAL, "invalid block");
  const size_t n = (size_t)d->n_elements * d->ncomp * c->s->ngll * c->s->ngll;
  dom_state* st = &c->st[di];
  from_float(st->ul, ul, n);
  from_float(st->fl, fl, n);
  const int rc = run_block(c->s, d, &d->blocks[bi], st->ul, st->fl);
  to_float(fl, st->fl, n);
  return rc;
}

int oracle_gather(oracle_ctx* c, int32_t di, const float* g, float* l) {
  if (!c || di < 0 || di >= c->s->n_domains) return fail(INSEM_HIP_EINVAL, "invalid domain");
  const insem_hip_domain* d = &c->s->domains[di];
  const size_t n = dom_len(d), nl = (size_t)d->n_elements * d->ncomp * c->s->ngll * c->s->ngll;
  dom_state* st = &c->st[di];
  real* x = (real*)calloc(n + 1, sizeof(real));
  from_float(x, g, n);
  gather(d, c->s->ngll * c->s->ngll, x, st->ul);
  free(x);
  to_float(l, st->ul, nl);
  return INSEM_HIP_OK;
}

int oracle_assemble(oracle_ctx* c, int32_t di, const float* l, float* g) {
  if (!c || di < 0 || di >= c->s->n_domains) return fail(INSEM_HIP_EINVAL, "invalid domain");
  const insem_hip_domain* d = &c->s->domains[di];
  const size_t n = dom_len(d), nl = (size_t)d->n_elements * d->ncomp * c->s->ngll * c->s->ngll;
  dom_state* st = &c->st[di];
  real* x = (real*)calloc(n + 1, sizeof(real));
  from_float(st->fl, l, nl);
  assemble(d, c->s->ngll * c->s->ngll, st->fl, x);
  to_float(g, x, n);
  free(x);
  return INSEM_HIP_OK;
}
