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
 * and the reverse mode of the fixed-dt Rusanov WENO-JS5 step for the
 * adjoint (timestepping.py:155-214, matrix-free; see the block before
 * oracle_num_threads).
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

/* ---------------------------------------------------------------------------
 * Adjoint of the fixed-dt SSP-RK3 loop for periodic Burgers with the Rusanov
 * flux and WENO-JS5 (the gradient of J = 0.5 ||u(T) - u_target||^2 dx with
 * respect to u(0)). The reference differentiates one step with jax.jacfwd,
 * a dense Jacobian whose transpose multiplies the cotangent
 * (timestepping.py:155-214, driven by drivers/burgers-adjoint.py:186-266);
 * this applies the same transpose matrix-free, in reverse through the
 * periodic fill, the reconstruction, the flux, the difference and the three
 * stages, with the forward pass checkpointed every `every` steps and each
 * window recomputed. Derivatives at the non-differentiable points follow
 * JAX's rules: d|x|/dx = sign(x) (0 at 0), and max passes half of its
 * cotangent to each operand at a tie.
 */

/* rows t00, t01, t10, t11, t20, t21 and h0, h1, h2 of side() as linear forms
 * in the stencil (vm2, vm1, v0, vp1, vp2) */
static const double kTForm[6][5] = {{1.0, -2.0, 1.0, 0.0, 0.0},
                                    {1.0, -4.0, 3.0, 0.0, 0.0},
                                    {0.0, 1.0, -2.0, 1.0, 0.0},
                                    {0.0, -1.0, 0.0, 1.0, 0.0},
                                    {0.0, 0.0, 1.0, -2.0, 1.0},
                                    {0.0, 0.0, 3.0, -4.0, 1.0}};
static const double kHForm[3][5] = {{2.0 / 6.0, -7.0 / 6.0, 11.0 / 6.0, 0.0, 0.0},
                                    {0.0, -1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0, 0.0},
                                    {0.0, 0.0, 2.0 / 6.0, 5.0 / 6.0, -1.0 / 6.0}};

/* side(v) and its gradient G[k] = d side / d v[k]: with e_k = eps + beta_k,
 * w_k the normalised weights and r the value, d r / d beta_k =
 * -2 w_k (h_k - r) / e_k, d beta_k / d t_kj = 2 a_j t_kj, and r also moves
 * with the candidates h_k through w_k. */
static double side_grad(const double v[5], double eps, double G[5]) {
  const double aj[2] = {13.0 / 12.0, 1.0 / 4.0};
  const double dk[3] = {1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0};
  double t[6], h[3], e[3], al[3];
  for (int q = 0; q < 6; ++q) {
    t[q] = 0.0;
    for (int k = 0; k < 5; ++k) t[q] += kTForm[q][k] * v[k];
  }
  for (int q = 0; q < 3; ++q) {
    h[q] = 0.0;
    for (int k = 0; k < 5; ++k) h[q] += kHForm[q][k] * v[k];
    const double beta = aj[0] * t[2 * q] * t[2 * q] + aj[1] * t[2 * q + 1] * t[2 * q + 1];
    e[q] = eps + beta;
    al[q] = dk[q] / (e[q] * e[q]);
  }
  const double s = al[0] + al[1] + al[2];
  const double r = side(v[0], v[1], v[2], v[3], v[4], eps);
  for (int k = 0; k < 5; ++k) G[k] = 0.0;
  for (int q = 0; q < 3; ++q) {
    const double w = al[q] / s;
    const double dbeta = -2.0 * w * (h[q] - r) / e[q];
    for (int k = 0; k < 5; ++k) {
      G[k] += w * kHForm[q][k];
      G[k] += dbeta * (2.0 * aj[0] * t[2 * q] * kTForm[2 * q][k] +
                       2.0 * aj[1] * t[2 * q + 1] * kTForm[2 * q + 1][k]);
    }
  }
  return r;
}

static inline double sign_of(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

/* workspace of one operator transpose over m cells */
typedef struct {
  double *uf, *ur, *ul, *GR, *GL, *gur, *gul, *sR, *sL, *guf;
} VjpWork;

static int vjp_work_alloc(VjpWork *w, int64_t m) {
  const size_t b = (size_t)m * sizeof(double);
  w->uf = malloc(b);
  w->ur = malloc(b);
  w->ul = malloc(b);
  w->GR = malloc(5 * b);
  w->GL = malloc(5 * b);
  w->gur = malloc(b);
  w->gul = malloc(b);
  w->sR = malloc(b + sizeof(double));
  w->sL = malloc(b + sizeof(double));
  w->guf = malloc(b);
  return w->uf && w->ur && w->ul && w->GR && w->GL && w->gur && w->gul && w->sR &&
         w->sL && w->guf;
}

static void vjp_work_free(VjpWork *w) {
  free(w->uf);
  free(w->ur);
  free(w->ul);
  free(w->GR);
  free(w->GL);
  free(w->gur);
  free(w->gul);
  free(w->sR);
  free(w->sL);
  free(w->guf);
}

/* out = J_L(u)^T mu for L = apply_operator (Rusanov, WENO-JS5, periodic):
 * every entry of mu is a cotangent of L, ghosts included; out's ghost
 * entries are zero (the periodic fill overwrites them before they are read) */
static void operator_vjp(const double *u, const double *mu, double *out, int64_t m,
                         int g, double dx, double eps, VjpWork *w) {
  const int64_t n = m - 2 * (int64_t)g;
  memcpy(w->uf, u, (size_t)m * sizeof(double));
  fill_periodic(w->uf, m, g);
  const double *uf = w->uf;
  /* the two states of every cell and their stencil gradients */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    double v[5], vm[5];
    for (int k = 0; k < 5; ++k) {
      v[k] = at(uf, i - 2 + k, m);
      vm[k] = at(uf, i + 2 - k, m);
    }
    w->ur[i] = side_grad(v, eps, &w->GR[5 * i]);
    w->ul[i] = side_grad(vm, eps, &w->GL[5 * i]);
  }
  /* face j (cells j - 1, j) feeds L[j - 1] with -1/dx and L[j] with +1/dx */
  w->gur[m - 1] = 0.0;
  w->gul[0] = 0.0;
  w->sR[0] = w->sL[0] = 0.0;
  w->sR[m] = w->sL[m] = 0.0;
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j < m; ++j) {
    const double gF = (mu[j] - mu[j - 1]) / dx;
    const double urL = w->ur[j - 1], ulR = w->ul[j];
    const double aR = fabs(uf[j]), aL = fabs(uf[j - 1]);
    const double a = aR > aL ? aR : aL;
    /* F = 0.5 (ulR^2 / 2 + urL^2 / 2) - 0.5 a (ulR - urL) */
    w->gur[j - 1] = gF * (0.5 * urL + 0.5 * a);
    w->gul[j] = gF * (0.5 * ulR - 0.5 * a);
    const double ga = gF * (-0.5 * (ulR - urL));
    const double share = aR > aL ? 1.0 : (aR < aL ? 0.0 : 0.5);
    w->sR[j] = ga * share * sign_of(uf[j]);
    w->sL[j] = ga * (1.0 - share) * sign_of(uf[j - 1]);
  }
  /* gather: cell c is entry k of the right stencil of cell c + 2 - k and of
   * the mirrored left stencil of cell c - 2 + k */
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < m; ++c) {
    double acc = w->sR[c] + w->sL[c + 1];
    for (int k = 0; k < 5; ++k) {
      const int64_t ir = c + 2 - k, il = c - 2 + k;
      if (ir >= 0 && ir < m) acc += w->gur[ir] * w->GR[5 * ir + k];
      if (il >= 0 && il < m) acc += w->gul[il] * w->GL[5 * il + k];
    }
    w->guf[c] = acc;
  }
  /* transpose of the periodic fill: ghost k copies cell n + k, ghost
   * m - g + k copies cell g + k */
  for (int64_t i = 0; i < m; ++i) out[i] = (i >= g && i < m - g) ? w->guf[i] : 0.0;
  for (int k = 0; k < g; ++k) {
    out[n + k] += w->guf[k];
    out[g + k] += w->guf[m - g + k];
  }
}

void oracle_operator_vjp(const double *u, const double *mu, double *out, int64_t m,
                         int g, double dx, double eps) {
  VjpWork w;
  if (vjp_work_alloc(&w, m)) operator_vjp(u, mu, out, m, g, dx, eps, &w);
  vjp_work_free(&w);
}

/* buffers of one fixed-dt step (oracle_ssprk33_run's sequence) */
typedef struct {
  double *L, *k1, *k2, *uf, *F;
} StepWork;

static int step_work_alloc(StepWork *s, int64_t m) {
  const size_t b = (size_t)m * sizeof(double);
  s->L = malloc(b);
  s->k1 = malloc(b);
  s->k2 = malloc(b);
  s->uf = malloc(b);
  s->F = malloc(b + sizeof(double));
  return s->L && s->k1 && s->k2 && s->uf && s->F;
}

static void step_work_free(StepWork *s) {
  free(s->L);
  free(s->k1);
  free(s->k2);
  free(s->uf);
  free(s->F);
}

/* the stages k1, k2 of a Rusanov WENO-JS5 step from u, and with unext the
 * step's output, in the reference's expression order */
static void rusanov_step(const double *u, double *unext, int64_t m, int g, double dx,
                         double eps, double dt, StepWork *s) {
  oracle_apply_operator(ORACLE_RUSANOV, u, NULL, s->L, s->uf, s->F, m, g, dx, eps);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) s->k1[i] = u[i] + dt * s->L[i];
  oracle_apply_operator(ORACLE_RUSANOV, s->k1, NULL, s->L, s->uf, s->F, m, g, dx, eps);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i)
    s->k2[i] = 3.0 / 4.0 * u[i] + 1.0 / 4.0 * (s->k1[i] + dt * s->L[i]);
  if (unext == NULL) return;
  oracle_apply_operator(ORACLE_RUSANOV, s->k2, NULL, s->L, s->uf, s->F, m, g, dx, eps);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i)
    unext[i] = 1.0 / 3.0 * u[i] + 2.0 / 3.0 * (s->k2[i] + dt * s->L[i]);
}

/* out = (d step / d u)^T lam for the step from u: with the stage states
 * recomputed, lk2 = 2/3 (lam + dt J_L(k2)^T lam), lk1 = 1/4 (lk2 + dt
 * J_L(k1)^T lk2), out = 1/3 lam + 3/4 lk2 + lk1 + dt J_L(u)^T lk1
 * (the transpose of timestepping.py:312-320); a, b are scratch of m */
static void step_vjp(const double *u, const double *lam, double *out, int64_t m, int g,
                     double dx, double eps, double dt, StepWork *s, VjpWork *w,
                     double *a, double *b) {
  rusanov_step(u, NULL, m, g, dx, eps, dt, s);
  operator_vjp(s->k2, lam, a, m, g, dx, eps, w);
  for (int64_t i = 0; i < m; ++i) a[i] = 2.0 / 3.0 * (lam[i] + dt * a[i]); /* lk2 */
  operator_vjp(s->k1, a, b, m, g, dx, eps, w);
  for (int64_t i = 0; i < m; ++i) b[i] = 1.0 / 4.0 * (a[i] + dt * b[i]); /* lk1 */
  operator_vjp(u, b, out, m, g, dx, eps, w);
  for (int64_t i = 0; i < m; ++i)
    out[i] = (1.0 / 3.0 * lam[i] + 3.0 / 4.0 * a[i]) + (b[i] + dt * out[i]);
}

void oracle_ssprk33_step_vjp(const double *u, const double *lam, double *out, int64_t m,
                             int g, double dx, double eps, double dt) {
  StepWork s = {0};
  VjpWork w = {0};
  double *a = malloc((size_t)m * sizeof(double));
  double *b = malloc((size_t)m * sizeof(double));
  const int ok_s = step_work_alloc(&s, m), ok_w = vjp_work_alloc(&w, m);
  if (ok_s && ok_w && a && b)
    step_vjp(u, lam, out, m, g, dx, eps, dt, &s, &w, a, b);
  step_work_free(&s);
  vjp_work_free(&w);
  free(a);
  free(b);
}

/*
 * The gradient of J = 0.5 sum_interior (u(T) - target)^2 dx with respect to
 * u0 after nsteps steps of a fixed dt: the forward pass keeps u every `every`
 * steps, then each window, last first, is recomputed from its checkpoint and
 * the cotangent (u(T) - target) dx is carried back through its steps. On
 * return grad (ghosts zero), u_final and *J_out. Returns 0, or -1 when
 * memory runs out.
 */
int oracle_adjoint_rusanov(const double *u0, const double *target, double *grad,
                           double *u_final, double *J_out, int64_t n, int g, double dx,
                           double eps, double dt, long long nsteps, long long every) {
  const int64_t m = n + 2 * (int64_t)g;
  const size_t b = (size_t)m * sizeof(double);
  if (every < 1) every = 1;
  const long long ncp = nsteps > 0 ? (nsteps + every - 1) / every : 0;
  double *cps = malloc((ncp > 0 ? (size_t)ncp : 1) * b);
  double *win = malloc((size_t)every * b);
  double *x = malloc(b), *y = malloc(b), *lam = malloc(b), *tmp = malloc(b);
  double *va = malloc(b), *vb = malloc(b);
  StepWork s = {0};
  VjpWork w = {0};
  int status = 0;
  const int ok_s = step_work_alloc(&s, m), ok_w = vjp_work_alloc(&w, m);
  if (!(ok_s && ok_w && cps && win && x && y && lam && tmp && va && vb)) {
    status = -1;
    goto done;
  }
  memcpy(x, u0, b);
  for (long long st = 0; st < nsteps; ++st) {
    if (st % every == 0) memcpy(cps + (size_t)(st / every) * (size_t)m, x, b);
    rusanov_step(x, y, m, g, dx, eps, dt, &s);
    double *t = x;
    x = y;
    y = t;
  }
  memcpy(u_final, x, b);
  double J = 0.0;
  for (int64_t i = 0; i < m; ++i) {
    if (i >= g && i < m - g) {
      const double d = x[i] - target[i];
      J += d * d;
      lam[i] = d * dx;
    } else {
      lam[i] = 0.0;
    }
  }
  *J_out = 0.5 * J * dx;
  for (long long wi = ncp - 1; wi >= 0; --wi) {
    const long long s0 = wi * every;
    const long long s1 = s0 + every < nsteps ? s0 + every : nsteps;
    memcpy(win, cps + (size_t)wi * (size_t)m, b);
    for (long long st = s0 + 1; st < s1; ++st)
      rusanov_step(win + (size_t)(st - 1 - s0) * (size_t)m,
                   win + (size_t)(st - s0) * (size_t)m, m, g, dx, eps, dt, &s);
    for (long long st = s1 - 1; st >= s0; --st) {
      step_vjp(win + (size_t)(st - s0) * (size_t)m, lam, tmp, m, g, dx, eps, dt, &s, &w,
               va, vb);
      double *t = lam;
      lam = tmp;
      tmp = t;
    }
  }
  memcpy(grad, lam, b);
done:
  step_work_free(&s);
  vjp_work_free(&w);
  free(cps);
  free(win);
  free(x);
  free(y);
  free(lam);
  free(tmp);
  free(va);
  free(vb);
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
