/*
 * CPU oracle, C restatement of pyshocks' WENO-JS5 (and WENO-JS3, ES-WENO3) /
 * Rusanov / Godunov / Engquist-Osher / Lax-Friedrichs / SSP-RK3 hot path
 * (FP64), parallelised over cells with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: it is loaded by tests/ (checked against the
 * NumPy restatement oracle/weno_np.py) and timed as the CPU baseline by
 * bench.py. Nothing in the product library links or calls it.
 *
 * Per step it follows the reference's sequence and expression order:
 *   apply_boundary(Periodic)          scalar.py:529-540
 *   reconstruct(WENOJS53) both sides  weno.py:114-157,247-256,
 *                                     reconstruction.py:351-377
 *   numerical flux                    scalar.py:91-132 (upwind),
 *                                     scalar.py:192-249 (Rusanov),
 *                                     scalar.py:258-278 (Lax-Friedrichs),
 *                                     scalar.py:287-322 (Engquist-Osher,
 *                                     omega = 0),
 *                                     advection/schemes.py:100-114
 *   operator -(F[1:]-F[:-1])/dx       schemes.py:339-346,
 *                                     advection/schemes.py:62-73
 *   SSP-RK3 combination               timestepping.py:312-320
 *   step control                      timestepping.py:89-152 with
 *                                     predict_timestep of
 *                                     burgers/schemes.py:42-49,121-127 and
 *                                     advection/schemes.py:51-59
 * Arrays outside [0, m) are zero, as in the reference's 'same' convolution.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_RUSANOV 0
#define ORACLE_BURGERS_UPWIND 1
#define ORACLE_ADVECTION_UPWIND 2
#define ORACLE_ENGQUIST_OSHER 3
#define ORACLE_LAX_FRIEDRICHS 4

static inline double at(const double *u, int64_t i, int64_t m) {
  return (i >= 0 && i < m) ? u[i] : 0.0;
}

/* one side, reference expression order; v2..vm2 = u[i+2] .. u[i-2] */
static inline double side(double vm2, double vm1, double v0, double vp1,
                          double vp2, double eps) {
  const double a0 = 13.0 / 12.0, a1 = 1.0 / 4.0;
  const double t00 = (v0 - 2.0 * vm1) + vm2;
  const double t01 = (3.0 * v0 - 4.0 * vm1) + vm2;
  const double t10 = (vp1 - 2.0 * v0) + vm1;
  const double t11 = vp1 - vm1;
  const double t20 = (vp2 - 2.0 * vp1) + v0;
  const double t21 = (vp2 - 4.0 * vp1) + 3.0 * v0;
  const double b0 = a0 * (t00 * t00) + a1 * (t01 * t01);
  const double b1 = a0 * (t10 * t10) + a1 * (t11 * t11);
  const double b2 = a0 * (t20 * t20) + a1 * (t21 * t21);
  const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const double al0 = (1.0 / 10.0) / (e0 * e0);
  const double al1 = (6.0 / 10.0) / (e1 * e1);
  const double al2 = (3.0 / 10.0) / (e2 * e2);
  const double s = (al0 + al1) + al2;
  const double h0 = ((11.0 / 6.0) * v0 - (7.0 / 6.0) * vm1) + (2.0 / 6.0) * vm2;
  const double h1 = ((2.0 / 6.0) * vp1 + (5.0 / 6.0) * v0) - (1.0 / 6.0) * vm1;
  const double h2 = (-(1.0 / 6.0) * vp2 + (5.0 / 6.0) * vp1) + (2.0 / 6.0) * v0;
  return ((al0 / s) * h0 + (al1 / s) * h1) + (al2 / s) * h2;
}

static inline double recon_right(const double *u, int64_t i, int64_t m,
                                 double eps) {
  return side(at(u, i - 2, m), at(u, i - 1, m), u[i], at(u, i + 1, m),
              at(u, i + 2, m), eps);
}

static inline double recon_left(const double *u, int64_t i, int64_t m,
                                double eps) {
  return side(at(u, i + 2, m), at(u, i + 1, m), u[i], at(u, i - 1, m),
              at(u, i - 2, m), eps);
}

/* third-order side (weno.py:166-203 with the weights of weno.py:247-256, or
 * the ES-WENO weights of weno.py:285-296 with this tau); (a, b, c) =
 * (u[i-1], u[i], u[i+1]) for the right state, mirrored for the left */
static inline double side3(double a, double b, double c, double eps, int es,
                           double tau) {
  const double b0 = (b - a) * (b - a), b1 = (c - b) * (c - b);
  const double h0 = (3.0 / 2.0) * b + (-1.0 / 2.0) * a;
  const double h1 = (1.0 / 2.0) * c + (1.0 / 2.0) * b;
  double a0, a1;
  if (es) {
    a0 = (1.0 / 3.0) * (1.0 + tau / (eps + b0));
    a1 = (2.0 / 3.0) * (1.0 + tau / (eps + b1));
  } else {
    a0 = (1.0 / 3.0) / ((eps + b0) * (eps + b0));
    a1 = (2.0 / 3.0) / ((eps + b1) * (eps + b1));
  }
  const double s = a0 + a1;
  return (a0 / s) * h0 + (a1 / s) * h1;
}

/* rec 1 (WENO-JS3) or 2 (ES-WENO3); tau is zero at the array ends */
static inline double recon3(const double *u, int64_t i, int64_t m, double eps,
                            int rec, int right) {
  const double a = at(u, i - 1, m), b = u[i], c = at(u, i + 1, m);
  const int es = rec == 2;
  double tau = 0.0;
  if (es && i > 0 && i < m - 1) {
    const double t = (c - 2.0 * b) + a;
    tau = t * t;
  }
  return right ? side3(a, b, c, eps, es, tau) : side3(c, b, a, eps, es, tau);
}

/* flux at face j (between cells j - 1 and j), 1 <= j <= m - 1; amax is
 * the global speed of Lax-Friedrichs; rec 0 WENO-JS5, 1 WENO-JS3,
 * 2 ES-WENO3 */
static inline double face_flux(int flux, int rec, const double *u,
                               const double *vel, int64_t j, int64_t m,
                               double eps, double amax) {
  const double urL = rec ? recon3(u, j - 1, m, eps, rec, 1)
                         : recon_right(u, j - 1, m, eps);
  const double ulR = rec ? recon3(u, j, m, eps, rec, 0) : recon_left(u, j, m, eps);
  if (flux == ORACLE_RUSANOV || flux == ORACLE_LAX_FRIEDRICHS) {
    const double fl = (ulR * ulR) / 2.0;
    const double fr = (urL * urL) / 2.0;
    const double a = flux == ORACLE_RUSANOV ? fmax(fabs(u[j]), fabs(u[j - 1])) : amax;
    return 0.5 * (fl + fr) - (0.5 * a) * (ulR - urL);
  }
  if (flux == ORACLE_ENGQUIST_OSHER) {
    /* f(max(ur, 0)) + f(min(ul, 0)) - f(0), f(0) = 0 */
    const double p = fmax(urL, 0.0), q = fmin(ulR, 0.0);
    return (p * p) / 2.0 + (q * q) / 2.0;
  }
  if (flux == ORACLE_BURGERS_UPWIND) {
    const double aavg = (urL + ulR) / 2.0;
    return aavg > 0.0 ? (urL * urL) / 2.0 : (ulR * ulR) / 2.0;
  }
  {
    const double arL = rec ? recon3(vel, j - 1, m, eps, rec, 1)
                           : recon_right(vel, j - 1, m, eps);
    const double alR = rec ? recon3(vel, j, m, eps, rec, 0)
                           : recon_left(vel, j, m, eps);
    const double aavg = (arL + alR) / 2.0;
    return aavg > 0.0 ? urL : ulR;
  }
}

static void fill_periodic(double *u, int64_t m, int g) {
  const int64_t n = m - 2 * (int64_t)g;
  for (int k = 0; k < g; ++k) {
    u[m - g + k] = u[g + k];
    u[k] = u[n + k];
  }
}

/* L = apply_operator(u) into out (m entries); uf is scratch (m). flux
 * carries the reconstruction in its bits 8 and up (0: WENO-JS5). */
void oracle_apply_operator(int flux, const double *u, const double *vel,
                           double *out, double *uf, double *F, int64_t m,
                           int g, double dx, double eps) {
  const int rec = flux >> 8;
  flux &= 0xff;
  memcpy(uf, u, (size_t)m * sizeof(double));
  fill_periodic(uf, m, g);
  F[0] = 0.0;
  F[m] = 0.0;
  double amax = 0.0;
  if (flux == ORACLE_LAX_FRIEDRICHS) {
    /* max|a| over the whole ghost-filled array, a = u (scalar.py:276) */
#pragma omp parallel for reduction(max : amax) schedule(static)
    for (int64_t i = 0; i < m; ++i) {
      const double v = fabs(uf[i]);
      if (v > amax) amax = v;
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j < m; ++j)
    F[j] = face_flux(flux, rec, uf, vel, j, m, eps, amax);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    const double d = F[i + 1] - F[i];
    out[i] = flux == ORACLE_ADVECTION_UPWIND ? (-vel[i] * d) / dx : -d / dx;
  }
}

static double max_abs_interior(const double *u, int64_t m, int g) {
  double smax = 0.0;
#pragma omp parallel for reduction(max : smax) schedule(static)
  for (int64_t i = g; i < m - g; ++i) {
    const double v = fabs(u[i]);
    if (v > smax) smax = v;
  }
  return smax;
}

/*
 * Run the reference step loop. dt_mode 0: dt = theta * (cfl_num / smax),
 * 1: dt = fixed_dt. maxit < 0: no limit. tfinal as in the reference
 * (DBL_MAX when none). On return u holds the final state, *t_out and *m_out
 * the final time and step count, dts (if not NULL, capacity max_dts) the
 * time steps. Returns 0, or -4 for a non-finite time step.
 */
int oracle_ssprk33_run(int flux, double *u, const double *vel, int64_t n,
                       int g, double dx, double eps, int dt_mode, double theta,
                       double cfl_num, double fixed_dt, long long maxit,
                       double tstart, double tfinal, double *t_out,
                       long long *m_out, double *dts, long long max_dts) {
  const int64_t m = n + 2 * (int64_t)g;
  double *L = (double *)malloc((size_t)m * sizeof(double));
  double *k1 = (double *)malloc((size_t)m * sizeof(double));
  double *k2 = (double *)malloc((size_t)m * sizeof(double));
  double *uf = (double *)malloc((size_t)m * sizeof(double));
  double *F = (double *)malloc((size_t)(m + 1) * sizeof(double));
  double amax = 0.0;
  const int kind = flux & 0xff;  /* flux with the reconstruction bits */
  if (kind == ORACLE_ADVECTION_UPWIND) amax = max_abs_interior(vel, m, g);

  double t = tstart;
  long long it = 0;
  int status = 0;
  for (;;) {
    if (t >= tfinal) break;
    if (maxit >= 0 && it >= maxit) break;

    double dt;
    if (dt_mode == 1) {
      dt = fixed_dt;
    } else {
      const double smax =
          kind == ORACLE_ADVECTION_UPWIND ? amax : max_abs_interior(u, m, g);
      dt = theta * (cfl_num / smax);
    }
    if (!isinf(tfinal)) {
      const double dt_min = tfinal - t;
      dt = (dt < dt_min ? dt : dt_min) + 1.0e-15;
    }
    if (!isfinite(dt)) {
      status = -4;
      break;
    }

    oracle_apply_operator(flux, u, vel, L, uf, F, m, g, dx, eps);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) k1[i] = u[i] + dt * L[i];
    oracle_apply_operator(flux, k1, vel, L, uf, F, m, g, dx, eps);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
      k2[i] = 3.0 / 4.0 * u[i] + 1.0 / 4.0 * (k1[i] + dt * L[i]);
    oracle_apply_operator(flux, k2, vel, L, uf, F, m, g, dx, eps);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
      u[i] = 1.0 / 3.0 * u[i] + 2.0 / 3.0 * (k2[i] + dt * L[i]);

    if (dts != NULL && it < max_dts) dts[it] = dt;
    it += 1;
    t += dt;
  }

  free(L);
  free(k1);
  free(k2);
  free(uf);
  free(F);
  *t_out = t;
  *m_out = it;
  return status;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}
