/*
 * pyshocks_amd: C ABI of the MI355X (gfx950) WENO-JS5 / Rusanov / Godunov /
 * SSP-RK3 library.
 *
 * Conventions
 * -----------
 * - Every function returns an int status: PSA_OK (0) or a negative PSA_ERR_*
 *   code; psa_get_last_error() returns the thread-local message of the last
 *   failure.
 * - Arrays are contiguous device pointers of the given dtype (PSA_FLOAT64 or
 *   PSA_FLOAT32). A grid of n interior cells and g = 3 ghost cells per side
 *   has arrays of m = n + 2 g cells; face arrays have m + 1 entries.
 * - The caller owns every buffer. No function keeps a pointer past its return
 *   or allocates device memory (psa_ssprk33_solve_static uses a static
 *   device array instead of caller buffers); all work is enqueued
 *   asynchronously on the given hipStream_t (passed as void*; NULL is the
 *   default stream).
 * - Reference citations are relative to the pyshocks 2025.10 source tree
 *   (src/pyshocks/...), which this library replaces for the hot path.
 */
#ifndef PYSHOCKS_AMD_H_
#define PYSHOCKS_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes ----------------------------------------------------------- */
#define PSA_OK 0
#define PSA_ERR_INVALID_ARGUMENT -1 /* bad shape / pointer / parameter */
#define PSA_ERR_NOT_IMPLEMENTED -2  /* unsupported dtype / flux / mode */
#define PSA_ERR_HIP -3              /* HIP runtime error (launch, device) */
#define PSA_ERR_NONFINITE_DT -4     /* timestepping.py:144-145 */
#define PSA_ERR_WORKSPACE -5        /* workspace too small */

/* dtypes ------------------------------------------------------------------ */
#define PSA_FLOAT64 0
#define PSA_FLOAT32 1

/* numerical fluxes ---------------------------------------------------------
 * PSA_FLUX_RUSANOV          burgers.Rusanov  -> scalar_flux_rusanov
 *                           (burgers/schemes.py:99-118, scalar.py:192-249)
 * PSA_FLUX_BURGERS_UPWIND   burgers.Godunov  -> scalar_flux_upwind
 *                           (burgers/schemes.py:75-89, scalar.py:91-132)
 * PSA_FLUX_ADVECTION_UPWIND advection.Godunov -> upwind_flux
 *                           (advection/schemes.py:100-129)
 * PSA_FLUX_ENGQUIST_OSHER   burgers.EngquistOsher -> scalar_flux_engquist_osher
 *                           with omega = 0 (burgers/schemes.py:162-196,
 *                           scalar.py:287-322)
 * PSA_FLUX_LAX_FRIEDRICHS   burgers.LaxFriedrichs -> scalar_flux_lax_friedrichs
 *                           (burgers/schemes.py:136-153, scalar.py:258-278):
 *                           the Rusanov flux with the global speed max|a|
 *                           in place of the local one. Without a control
 *                           block that speed is a device scalar the caller
 *                           passes (see psa_numerical_flux); with one, the
 *                           stage kernels reduce it themselves.          */
#define PSA_FLUX_RUSANOV 0
#define PSA_FLUX_BURGERS_UPWIND 1
#define PSA_FLUX_ADVECTION_UPWIND 2
#define PSA_FLUX_ENGQUIST_OSHER 3
#define PSA_FLUX_LAX_FRIEDRICHS 4
/* ES-WENO for Burgers (burgers/schemes.py:230-256): the upwind flux of the
 * ES-WENO3 reconstruction plus the dissipative flux of its first weight
 * (psa_esweno_dissipation); fused stages and steps only, with
 * rec = PSA_REC_ESWENO32 and es_delta set. */
#define PSA_FLUX_ESWENO 5

/* reconstructions ----------------------------------------------------------
 * PSA_REC_WENOJS53  reconstruction.WENOJS53 (reconstruction.py:336-348,
 *                   weno.py:206-256), nghosts >= 3
 * PSA_REC_WENOJS32  reconstruction.WENOJS32 (reconstruction.py:320-333,
 *                   weno.py:166-203), nghosts >= 2
 * PSA_REC_ESWENO32  reconstruction.ESWENO32: the WENO-JS3 stencils with the
 *                   ES-WENO weights (reconstruction.py:387-434,
 *                   weno.py:285-296), nghosts >= 2                       */
#define PSA_REC_WENOJS53 0
#define PSA_REC_WENOJS32 1
#define PSA_REC_ESWENO32 2
/* PSA_REC_MUSCL     reconstruction.MUSCL, the flux-limited second-order
 *                   reconstruction (reconstruction.py:172-252) with the
 *                   limiter PSA_LIMITER_* of limiters.py:132-408,
 *                   nghosts >= 2                                         */
#define PSA_REC_MUSCL 3
/* PSA_REC_MUSCLS    reconstruction.MUSCLS, the slope-limited form
 *                   (reconstruction.py:255-296, limiters.py slope_limit):
 *                   ur, ul = f +- dx / 2 du with the limited slope du of
 *                   the limiter (not PSA_LIMITER_KOREN), nghosts >= 2     */
#define PSA_REC_MUSCLS 4

/* limiters (limiters.py:132-408); MINMOD takes theta in [1, 2] */
#define PSA_LIMITER_MINMOD 0
#define PSA_LIMITER_MC 1
#define PSA_LIMITER_SUPERBEE 2
#define PSA_LIMITER_VANALBADA1 3
#define PSA_LIMITER_VANALBADA2 4
#define PSA_LIMITER_VANLEER 5
#define PSA_LIMITER_KOREN 6
#define PSA_LIMITER_NONE 7

/* time step modes of the device-side step control */
#define PSA_DT_CFL 0   /* dt = theta * cfl_num / max|u_int| (or max|a_int|) */
#define PSA_DT_FIXED 1 /* dt = fixed_dt every step                           */

/* Library / error queries. */
int psa_version(void);
int psa_control_bytes(void); /* sizeof(psa_control) */
int psa_get_last_error(char *buf, size_t n);
int psa_device_count(int *count);

/* weno.py:114-140 weno_smoothness(weno_js_53_coefficients(), u):
 * beta (3, m), row-major, zero padding outside [0, m) as jnp.convolve('same'). */
int psa_weno_smoothness(int dtype, const void *u, void *beta, int64_t m,
                        void *stream);

/* weno.py:143-157 weno_interp: uhat (3, m). */
int psa_weno_interp(int dtype, const void *u, void *uhat, int64_t m,
                    void *stream);

/* weno.py:247-256 weno_js_weights (p = 2): omega (3, m). */
int psa_weno_js_weights(int dtype, const void *u, void *omega, int64_t m,
                        double eps, void *stream);

/* reconstruction.py:358-377 reconstruct(WENOJS53(eps), ...): (ul, ur), each
 * of m cells; ul is the value at x_{i-1/2}, ur at x_{i+1/2}. */
int psa_weno_reconstruct(int dtype, const void *f, void *ul, void *ur,
                         int64_t m, double eps, void *stream);

/* The same for any WENO PSA_REC_* (reconstruction.py:358-377,420-434): the
 * 'same' convolutions pad with zeros; ES-WENO's tau is zero for the first
 * and last cell (weno.py:293). */
int psa_reconstruct(int dtype, int rec, const void *f, void *ul, void *ur,
                    int64_t m, double eps, void *stream);

/* reconstruction.py:205-252 reconstruct(MUSCL(lm, atol), ...) of f as given:
 * the slope ratio of limiters.py:110-129 (0 where |u[i+1] - u[i]| < atol,
 * 0 for the first and last cell), ur = f + phi(r) / 2 (f[i+1] - f[i]),
 * ul = f - phi(1/r) / 2 (f[i] - f[i-1]), ur[m-1] = ul[0] = 0. */
int psa_muscl_reconstruct(int dtype, int limiter, double theta, double atol,
                          const void *f, void *ul, void *ur, int64_t m,
                          void *stream);

/* reconstruction.py:279-296 reconstruct(MUSCLS(lm), ...) of f as given on a
 * uniform grid of cell size dx: du = slope_limit(lm) of the one-sided slopes
 * (limiters.py:148-371; zero for the first and last cell), ur = f + dx/2 du,
 * ul = f - dx/2 du. */
int psa_muscls_reconstruct(int dtype, int limiter, const void *f, void *ul,
                           void *ur, int64_t m, double dx, void *stream);

/* scalar.py:529-540 apply_boundary(PeriodicBoundary): out = u with the
 * g ghost cells per side filled with their periodic images (out may be u). */
int psa_apply_boundary_periodic(int dtype, const void *u, void *out,
                                int64_t m, int g, void *stream);

/* scalar.py:441-457 (Dirichlet) and scalar.py:477-511 (Neumann) ghost
 * fills of one or both sides of u (out may be u; the interior is copied):
 * Dirichlet sets the g ghost cells of a side to the g device values given
 * for it (g_d(t, x) of the ghost cell centres, left cells 0 .. g-1, right
 * cells m-g .. m-1; NULL leaves that side); Neumann sets ghost cell g-1-k to
 * u[g+k] - (2k+1) dx gl and ghost cell m-1-k to u[m-2g+k] + (2k+1-2g) dx gr
 * for k < g, the reference's u[from] + side (x[from] - x[to]) g_n on a
 * uniform grid (sides: 1 left, 2 right, 3 both). */
int psa_apply_boundary_dirichlet(int dtype, const void *u, void *out,
                                 int64_t m, int g, const void *left,
                                 const void *right, void *stream);
int psa_apply_boundary_neumann(int dtype, const void *u, void *out, int64_t m,
                               int g, double dx, double gl, double gr,
                               int sides, void *stream);

/* The dissipative flux of the ES-WENO schemes (burgers/schemes.py:230-256,
 * advection/schemes.py:162-186) added to fnum (m + 1) in place: at face j
 * (cells j - 1, j) for 1 <= j <= m - 1, with w the ES-WENO weight of stencil
 * `weight` (0 Burgers, 1 advection) of each cell of u (weno.py:284-296: zero
 * padding at the array ends, tau zero in the end cells) and
 * dw = w[j] - w[j - 1]: fnum[j] += -(sqrt(dw^2 + delta^2) / 8 + dw / 8)
 * (u[j] - u[j - 1]). */
int psa_esweno_dissipation(int dtype, const void *u, void *fnum, int64_t m,
                           double eps, double delta, int weight, void *stream);

/* The ES-WENO scheme for linear advection (advection/schemes.py:135-186) on
 * u as given: fnum (m + 1), zero at faces 0 and m, at face j
 * 0.5 (fl[j] + fr[j - 1]) - 0.5 amax (ul[j] - ur[j - 1]) with (ul, ur) the
 * ES-WENO3 states of u and (fl, fr) those of f = velocity u, amax ONE device
 * value of the dtype (max|velocity| over the whole array), plus the
 * dissipative flux of the second ES-WENO weight of u
 * (psa_esweno_dissipation with weight 1). */
int psa_esweno_advection_flux(int dtype, const void *u, const void *velocity,
                              const void *amax, void *fnum, int64_t m, double eps,
                              double delta, void *stream);

/* schemes.py:466-473 and advection/schemes.py:62-73 from face fluxes:
 * out[i] = -(fnum[i + 1] - fnum[i]) / dx for every cell, times velocity[i]
 * when velocity (m) is not null. */
int psa_flux_divergence(int dtype, const void *fnum, const void *velocity,
                        void *out, int64_t m, double dx, void *stream);

/* schemes.py:322-336 numerical_flux on u as given (no boundary applied):
 * fnum (m + 1) with fnum[0] = fnum[m] = 0. velocity (m) is required for
 * PSA_FLUX_ADVECTION_UPWIND; for PSA_FLUX_LAX_FRIEDRICHS it points to ONE
 * device value of the dtype, the global speed max|a| (scalar.py:276, e.g.
 * from psa_max_abs_interior with g = 0); it is ignored otherwise. nu (m - 1)
 * is the artificial-viscosity factor grid.df ** (alpha - 1) of the Rusanov
 * and Lax-Friedrichs fluxes, or NULL for alpha == 1. */
int psa_numerical_flux(int dtype, int flux, const void *u,
                       const void *velocity, const void *nu, void *fnum,
                       int64_t m, double eps, void *stream);

/* psa_numerical_flux with the reconstruction rec (PSA_REC_*); an advection
 * velocity array is reconstructed the same way (advection/schemes.py:108). */
int psa_numerical_flux_rec(int dtype, int flux, int rec, const void *u,
                           const void *velocity, const void *nu, void *fnum,
                           int64_t m, double eps, void *stream);

/* schemes.py:339-346 (conservation law) and advection/schemes.py:62-73:
 * out (m) = -(F[1:] - F[:-1]) / dx (times -velocity for advection), with F
 * the numerical flux of the periodic ghost-filled u; every entry, ghosts
 * included, as the reference computes it. For PSA_FLUX_LAX_FRIEDRICHS,
 * velocity is the device scalar max|u| of the ghost-filled state (the
 * interior maximum, psa_max_abs_interior). */
int psa_apply_operator(int dtype, int flux, const void *u,
                       const void *velocity, const void *nu, void *out,
                       int64_t m, int g, double dx, double eps, void *stream);

/* psa_apply_operator with the reconstruction rec (PSA_REC_*); an advection
 * velocity array is reconstructed the same way. The MUSCL reconstructions
 * and PSA_FLUX_ESWENO take their parameters from a descriptor
 * (psa_ssprk33_stage, stage 0) and are refused here. */
int psa_apply_operator_rec(int dtype, int flux, int rec, const void *u,
                           const void *velocity, const void *nu, void *out,
                           int64_t m, int g, double dx, double eps,
                           void *stream);

/* max |u[g:m-g]| (burgers/schemes.py:42-49,121-127 and
 * advection/schemes.py:51-59 smax / amax) into the device scalar out. */
int psa_max_abs_interior(int dtype, const void *u, void *out, int64_t m,
                         int g, void *stream);

/* grid.py:445-506 _norm / norm over the interior slice [lo, hi) of u:
 * PSA_NORM_L1 sum |u| * weight, PSA_NORM_L2 sqrt(sum u^2 * weight),
 * PSA_NORM_INF max |u| (NaN propagates), PSA_NORM_TVD sum ||u[i+1]| - |u[i]||.
 * out is a device scalar of the dtype; the workspace holds
 * psa_norm_workspace_bytes() bytes. Deterministic (fixed reduction order). */
#define PSA_NORM_L1 1
#define PSA_NORM_L2 2
#define PSA_NORM_INF 3
#define PSA_NORM_TVD 4
int psa_norm_workspace_bytes(void);
int psa_interior_norm(int dtype, int p, const void *u, int64_t lo, int64_t hi,
                      double weight, void *out, void *workspace,
                      size_t workspace_bytes, void *stream);

/* Fused SSP-RK3 -------------------------------------------------------------
 * One HIP kernel per Runge-Kutta stage: WENO-JS5 smoothness, weights and
 * both reconstructions, the face flux and the stage combination of
 * timestepping.py:312-320 in one pass over an LDS tile with a 3-cell halo.
 * The periodic ghost fill is read by wrapped index; outputs are the interior
 * cells with their ghost images rewritten.
 *
 * The step control of timestepping.py:89-152 runs on the device in a
 * psa_control block: stage 1 decides whether the step is taken
 * (t < tfinal and m < maxit), predicts dt (PSA_DT_CFL: theta * cfl_num /
 * max|u_int| of the current state, clipped to tfinal - t, plus 1e-15),
 * stage 2 advances t and m, stage 3 reduces max|u_int| of the new state.
 * A non-finite dt stops the run with status PSA_ERR_NONFINITE_DT. */
#define PSA_UMAX_SLOTS 64      /* max|u| partials, one 128-byte line each */
#define PSA_UMAX_SLOT_STRIDE 16 /* in unsigned long long                    */
#define PSA_STEP_SLOT_STRIDE 4  /* whole-step partials, in unsigned long long */

typedef struct psa_control {
  double t;                     /* time of the current state               */
  double dt;                    /* dt of the last step taken               */
  double tfinal;                /* final time (finfo(dtype).max if none)   */
  double theta;                 /* Courant factor                          */
  double cfl_num;               /* 0.5 dx^(2-alpha), 0.5 dx or dx          */
  double fixed_dt;              /* dt for PSA_DT_FIXED                     */
  long long m;                  /* steps taken                             */
  long long maxit;              /* step limit, < 0 for none                */
  unsigned long long umax_bits; /* max|u_int| of the state the last step
                                   started from (bit pattern of the dtype) */
  int active;                   /* 1 while the step in flight is taken     */
  int status;                   /* PSA_OK or PSA_ERR_NONFINITE_DT          */
  int dt_mode;                  /* PSA_DT_CFL or PSA_DT_FIXED              */
  int speed_from_state;         /* 1: max|u| of each state (Burgers with
                                   PSA_DT_CFL, and Lax-Friedrichs)         */
  /* whole-step schedule (PSA_SCHEDULE_STEP): t, m and status of the state
   * that launch s of a psa_ssprk33_steps call starts from, in entry s % 2
   * (every workgroup of launch s decides step s from them; workgroup 0
   * records the state after the step in entry (s + 1) % 2) */
  double seq_t[2];
  long long seq_m[2];
  int seq_status[2];
  /* stage 3 reduces max|u_int| of the new state into word 0 of these slots
   * (a workgroup uses slot blockIdx % PSA_UMAX_SLOTS); stage 1 of the next
   * step combines them and stage 2 clears them. After a psa_ssprk33_steps
   * call on the whole-step schedule, word 0 of slot 0 holds the combined
   * value of the final state (the other word-0 entries zero). */
  unsigned long long umax_slot[PSA_UMAX_SLOTS][PSA_UMAX_SLOT_STRIDE];
  /* Lax-Friedrichs: stage s (1, 2) reduces max|u| of its output into word 0
   * of speed_slot[s - 1], which stage s + 1 combines as its flux speed;
   * stage 1 clears speed_slot[1] and stage 3 clears speed_slot[0]. Lines of
   * their own: no launch reads the lines it updates atomically. */
  unsigned long long speed_slot[2][PSA_UMAX_SLOTS][PSA_UMAX_SLOT_STRIDE];
  /* whole-step schedule: launch s reduces max|u_int| of the new state into
   * word 0 of step_slot[s % 3] (slot blockIdx % PSA_UMAX_SLOTS), launch
   * s + 1 combines them in every workgroup for its dt, and workgroup 0 of
   * launch s clears step_slot[(s + 1) % 3] for launch s + 1 (its last
   * reader was launch s - 1). Four slots share a 128-byte line: few lines
   * for every workgroup to read, few atomics per line. */
  unsigned long long step_slot[3][PSA_UMAX_SLOTS][PSA_STEP_SLOT_STRIDE];
} psa_control;

typedef struct psa_solver_desc {
  int dtype;                    /* PSA_FLOAT64 / PSA_FLOAT32               */
  int flux;                     /* PSA_FLUX_*                              */
  int64_t n;                    /* interior cells                          */
  int g;                        /* ghost cells per side (>= 3)             */
  int velocity_is_const;        /* advection: velocity == velocity_const;
                                   velocity may then be NULL               */
  double dx;                    /* cell size h                             */
  double eps;                   /* WENO-JS epsilon                         */
  double velocity_const;        /* advection constant velocity             */
  const void *velocity;         /* advection velocity (m) if not const;
                                   Lax-Friedrichs without a control block:
                                   the device scalar max|a| (one value)    */
  const void *nu;               /* Rusanov / Lax-Friedrichs df^(alpha-1)
                                   (m - 1) or NULL                         */
  int block;                    /* threads per workgroup (0 = tuned)       */
  int cells_per_thread;         /* cells per thread (0 = tuned)            */
  const float *mlp;             /* learned weights (PSA_MLP_NPARAM FP32) or
                                   NULL for WENO-JS                        */
  double mlp_tau;               /* shock threshold of the learned weights  */
  double mlp_sigma;             /* smooth threshold of the learned weights */
  int schedule;                 /* stages 1-3: PSA_SCHEDULE_*              */
  int rec;                      /* PSA_REC_* (0: WENO-JS5)                 */
  int limiter;                  /* PSA_REC_MUSCL: PSA_LIMITER_*            */
  int pad2_;
  double limiter_theta;         /* PSA_LIMITER_MINMOD: theta in [1, 2]     */
  double muscl_atol;            /* PSA_REC_MUSCL: slope cut-off (1e-12)    */
  int bc;                       /* PSA_BC_PERIODIC; PSA_BC_GIVEN: the
                                   operator (stage 0) takes the ghost cells
                                   of its input as given (a Dirichlet or
                                   Neumann fill done beforehand);
                                   PSA_BC_NEUMANN / PSA_BC_DIRICHLET: the
                                   whole-step kernel of psa_ssprk33_steps
                                   fills the ghost cells of every stage
                                   input itself (below)                    */
  int pad3_;
  double es_delta;              /* PSA_FLUX_ESWENO: delta of the
                                   dissipative flux                        */
  /* PSA_BC_NEUMANN (scalar.py:466-511 with a constant g on each side): the
   * ghost cell at distance e from the boundary (e = 0 innermost) gets the
   * interior cell at distance e minus (2e + 1) dx bc_left (bc_right on the
   * right), the reference's u[to] = u[from] + side (x[from] - x[to]) g.
   * PSA_BC_DIRICHLET (scalar.py:441-460 with a g that does not depend on
   * t): bc_values points to 2 g device values of the dtype, the left ghost
   * cells 0 .. g-1 and then the right ones m-g .. m-1. */
  double bc_left;
  double bc_right;
  const void *bc_values;
} psa_solver_desc;
#define PSA_BC_PERIODIC 0
#define PSA_BC_GIVEN 1
#define PSA_BC_NEUMANN 2
#define PSA_BC_DIRICHLET 3

/* Step schedules of psa_ssprk33_steps:
 * TILE: one kernel per stage, one workgroup per tile (8 words of HBM traffic
 *   per cell and step);
 * PIPELINE: the stage kernels on a persistent grid that copies its next tile
 *   into LDS while computing the current one (WENO-JS weights, no velocity
 *   array);
 * STEP: one kernel per step (a one-wave kernel before and after the call's
 *   steps); every workgroup decides the step from the control block, loads
 *   its tile of u^n with a 12-cell margin on each side, runs the three
 *   stages in LDS (recomputing the margin cells of the first two) and writes
 *   u^(n+1) (2 words per cell and step; any reconstruction, WENO-JS or
 *   learned weights, no velocity array, no Rusanov nu, not Lax-Friedrichs
 *   (its stages need the global max|u| of the previous stage), a control
 *   block, tiles 256 x 4 or 256 x 2). Steps ping-pong between u_out and the
 *   stage buffers.
 * RING: one launch for all the steps of the call; one workgroup holds a
 *   periodic state of n <= PSA_RING_MAX_CELLS cells in LDS (one cell per
 *   thread) and takes the steps with the arithmetic and the decisions of
 *   STEP, with no HBM traffic between steps (what STEP takes, periodic, no
 *   learned weights; the stage buffers are not used).
 * DEFAULT is RING where it applies, else STEP where it applies, else TILE.
 * All schedules compute the same values. */
#define PSA_SCHEDULE_DEFAULT 0
#define PSA_SCHEDULE_TILE 1
#define PSA_SCHEDULE_PIPELINE 2
#define PSA_SCHEDULE_STEP 3
#define PSA_SCHEDULE_RING 4
#define PSA_RING_MAX_CELLS 1024

/* The numerical flux of the descriptor's flux and reconstruction (any
 * PSA_REC_*, MUSCL's limiter from the descriptor) on u as given, fnum (m + 1)
 * with fnum[0] = fnum[m] = 0; m = n + 2 g. speed is the Lax-Friedrichs
 * device scalar max|a| (NULL otherwise); the advection velocity and nu come
 * from the descriptor. */
int psa_numerical_flux_desc(const psa_solver_desc *desc, const void *u,
                            const void *speed, void *fnum, void *stream);

/* Initialise the control block (device memory, sizeof(psa_control)) for a
 * run starting from state u at time tstart: stores t, tfinal, maxit, the dt
 * mode and, for PSA_DT_CFL, max|u_int| (Burgers) or max|velocity_int|
 * (advection). maxit < 0 means no limit. */
int psa_control_init(const psa_solver_desc *desc, const void *u,
                     void *control, double tstart, double tfinal,
                     long long maxit, int dt_mode, double theta,
                     double cfl_num, double fixed_dt, void *stream);

/* Enqueue nsteps steps of timestepping.step with SSP-RK3: u_in -> u_out
 * (u_out may equal u_in for an in-place update), stage buffers k1, k2 of m
 * cells. Steps after the control block says stop are no-ops, so a fixed
 * number of steps can be enqueued (or captured in a graph) and the control
 * block read back afterwards. */
int psa_ssprk33_steps(const psa_solver_desc *desc, const void *u_in,
                      void *u_out, void *k1, void *k2, void *control,
                      long long nsteps, void *stream);

/* One stage (1, 2 or 3) of timestepping.py:312-320: stage 1 out = uk + dt L(uk),
 * stage 2 out = 3/4 un + 1/4 (uk + dt L(uk)), stage 3 out = 1/3 un +
 * 2/3 (uk + dt L(uk)), with the step control of the control block, or, if
 * control is NULL, the device scalar dt. out may equal un but not uk.
 * Stage 0 is the operator itself, out = L(uk) on all m cells
 * (schemes.py:466-473, as psa_apply_operator), for any descriptor,
 * including one with learned weights; un, control and dt are unused.
 * Lax-Friedrichs without a control block (stage 0, or a device dt) takes
 * the flux speed max|uk| from desc->velocity. */
int psa_ssprk33_stage(const psa_solver_desc *desc, int stage, const void *un,
                      const void *uk, void *out, void *control,
                      const void *dt, void *stream);

/* timestepping.py:312-320 advance(SSPRK33, dt, t, u) for the bound operator
 * with the dt of a device scalar of the dtype: out = advance(u). Not for
 * Lax-Friedrichs (PSA_ERR_NOT_IMPLEMENTED: its stages need a control block
 * to carry the stage maxima). */
int psa_ssprk33_advance(const psa_solver_desc *desc, const void *u,
                        void *out, void *k1, void *k2, const void *dt,
                        void *stream);

/* Batched solver ------------------------------------------------------------
 * ninst independent periodic Burgers instances (Rusanov with alpha = 1,
 * Godunov or Engquist-Osher) of uneven size, each advanced from t = 0 to its own tfinal with
 * its own CFL time step (timestepping.py:89-152) in one persistent launch:
 * every workgroup takes the next instance from `order` (an atomic counter)
 * and runs its whole time loop; the launch is graph-capturable (the counter
 * is reset by a kernel). Instance i occupies cells [offsets[i],
 * offsets[i] + ncells[i] + 2 g) of the packed arrays u, k1, k2, and its
 * offset must be a multiple of 16 bytes. All arrays in the descriptor are
 * device pointers. Outputs per instance: final time, steps taken, status
 * (PSA_OK; PSA_ERR_NONFINITE_DT for a non-finite dt or one that no longer
 * advances t; PSA_ERR_INVALID_ARGUMENT when max_steps is reached). */
typedef struct psa_batch_desc {
  int dtype;                    /* PSA_FLOAT64 / PSA_FLOAT32               */
  int flux;                     /* PSA_FLUX_RUSANOV / _BURGERS_UPWIND /
                                   _ENGQUIST_OSHER                         */
  int g;                        /* ghost cells per side (>= 3; >= 2 for
                                   the other reconstructions)              */
  int grid;                     /* workgroups (0 = resident capacity)      */
  long long ninst;              /* instances                               */
  const long long *offsets;     /* [ninst] first cell of each instance     */
  const int *ncells;            /* [ninst] interior cells                  */
  const double *dx;             /* [ninst] cell size                       */
  const double *tfinal;         /* [ninst] final time                      */
  const int *order;             /* [ninst] processing order                */
  double eps;                   /* WENO-JS epsilon                         */
  double theta;                 /* Courant factor                          */
  long long max_steps;          /* step limit per instance                 */
  int block;                    /* threads per workgroup (0 = default)     */
  int cells_per_thread;         /* cells per thread (0 = default)          */
  const float *mlp;             /* learned weights or NULL (see
                                   psa_solver_desc)                        */
  double mlp_tau;               /* shock threshold of the learned weights  */
  double mlp_sigma;             /* smooth threshold of the learned weights */
  int time_f64;                 /* 0: t and dt in the state dtype, as
                                   timestepping.py:136-152; 1: in FP64
                                   (FP32 states then end exactly at tfinal) */
  int max_team;                 /* largest team size (teams only)          */
  /* Teams (all NULL/0: one workgroup per instance, in `order`). Instance i
   * is advanced by team[i] workgroups that split its tiles and meet at a
   * device-wide barrier after every stage; the queue is then the nslots
   * slots (slot_inst[s], slot_rank[s]) listing each instance's members
   * consecutively, ranks 0 .. team[i] - 1. team_state is caller-owned device
   * scratch of ninst * PSA_TEAM_STATE_BYTES bytes, cleared by the call.
   * Every team must fit the grid, which must fit the device at once; a
   * member waiting longer than team_wait_s seconds (0 = 60) at a barrier
   * stops its instance with status PSA_ERR_HIP. */
  const int *team;              /* [ninst] team sizes                      */
  const int *slot_inst;         /* [nslots] instance of each slot          */
  const int *slot_rank;         /* [nslots] member rank of each slot       */
  long long nslots;
  void *team_state;
  double team_wait_s;
  /* PSA_SCHEDULE_STEP (or DEFAULT): one pass per step over an instance's
   * tiles with the three stages of a tile in LDS (step_tile of the
   * whole-step kernel; one barrier per step, the state ping-pongs between
   * u and k1); PSA_SCHEDULE_TILE: one pass and one barrier per stage. Both
   * compute the same values. PSA_FLUX_LAX_FRIEDRICHS (WENO-JS5 weights,
   * no learned weights) always takes the stage passes: each stage's flux
   * speed is max|u| of its input over the instance, reduced after the
   * stage before. */
  int schedule;
  int rec;                      /* PSA_REC_* (0 = WENO-JS5); the others take
                                   g >= 2 and tiles 256 x 4 or 256 x 2, and
                                   no learned weights                      */
  int limiter;                  /* PSA_LIMITER_* of PSA_REC_MUSCL(S)      */
  int pad2_;
  double limiter_theta;         /* minmod's theta (PSA_REC_MUSCL)          */
  double muscl_atol;            /* slope-ratio cut-off (PSA_REC_MUSCL)     */
} psa_batch_desc;

#define PSA_TEAM_STATE_BYTES 64

int psa_batch_solve(const psa_batch_desc *desc, void *u, void *k1, void *k2,
                    unsigned int *counter, double *t_out, long long *steps_out,
                    int *status_out, void *stream);

/* Resources of the kernel psa_batch_solve launches for desc (its dtype,
 * flux, reconstruction, learned weights, schedule, block and
 * cells_per_thread; the buffers are not read): registers per lane, private
 * (spill) bytes per lane and static LDS bytes per workgroup. Launches
 * nothing; the tile autotuner uses it to skip candidates that spill. Part
 * of the batched front end, which the reference does not have. */
int psa_batch_kernel_info(const psa_batch_desc *desc, int *vgprs,
                          long long *private_bytes, long long *lds_bytes);

/* Workgroups of the kernel psa_batch_solve launches for desc that the
 * current device holds at once: the persistent grid when desc->grid is 0,
 * and the most workgroups that a team may span. The host sizes teams from
 * it. Launches nothing and reads no buffer. Part of the batched front end,
 * which the reference does not have. */
int psa_batch_resident_workgroups(const psa_batch_desc *desc, int *workgroups);

/* Discrete adjoint --------------------------------------------------------
 * Replaces the dense jax.jacfwd Jacobian of timestepping.py:155-215
 * (adjoint_step) with matrix-free vector-Jacobian products of the fused
 * step. Cotangents are arrays of m cells whose ghost entries are ignored on
 * input and zero on output. Supported: Rusanov (alpha = 1), Burgers upwind,
 * Engquist-Osher, Lax-Friedrichs (alpha = 1; its global speed max|u| is
 * differentiated too; single instances) and advection with a constant
 * velocity; dt is a device scalar held constant. Non-smooth points follow jax's conventions (max ties
 * split 0.5 / 0.5, sign(0) = 0, jnp.where differentiates the selected
 * branch). The three Burgers fluxes also take the learned reconstruction (desc->mlp,
 * no counterpart in the reference): the transposed Jacobian is that of the
 * forward solver's surrogate with the regimes held fixed at the forward
 * state: shock cells differentiate WENO-JS, smooth cells the ideal weights
 * (a linear map), learned cells the network with its weights rounded to
 * BF16 as the forward rounds them and its activations and inputs in the
 * solver's precision (the forward rounds those to BF16, a step function
 * whose derivative is zero almost everywhere). */

/* Reverse of one stage. stage 0: out = J_L(state)^T mu. stage 3:
 * out = 2/3 (mu + dt J_L(state)^T mu) with state = k2 and mu the step
 * cotangent; stage 2: out = 1/4 (mu + dt J_L(state)^T mu) with state = k1;
 * stage 1: out = 1/3 lam + 3/4 lam_k2 + mu + dt J_L(state)^T mu with
 * state = u. */
int psa_ssprk33_stage_vjp(const psa_solver_desc *desc, int stage,
                          const void *state, const void *mu, const void *lam,
                          const void *lam_k2, void *out, const void *dt,
                          void *stream);

/* lam_out = J(u)^T lam for one SSP-RK3 step from u (the per-step product of
 * adjoint_step); work holds 4 arrays of m cells (each rounded up to 256
 * bytes). */
int psa_ssprk33_step_vjp(const psa_solver_desc *desc, const void *u,
                         const void *lam, void *lam_out, void *work,
                         size_t work_bytes, const void *dt, void *stream);

/* Checkpointed forward solve and its reverse sweep: nsteps fixed-dt steps
 * from u0 storing the state every `every` steps in the workspace, then the
 * gradient grad_u0 = (d u_final / d u0)^T lam_final by recomputing each
 * window of `every` steps from its checkpoint (keeping every stage state)
 * and sweeping it backwards. Both calls use the same workspace of
 * psa_adjoint_workspace_bytes() bytes. With every >= nsteps (one window over
 * the horizon, about 3 every arrays) the forward pass keeps every stage state
 * itself and the reverse sweep recomputes nothing; the adjoint may then be
 * run any number of times on the workspace the forward pass left. */
int psa_adjoint_workspace_bytes(const psa_solver_desc *desc, long long nsteps,
                                int every, size_t *bytes);
int psa_ssprk33_forward_checkpointed(const psa_solver_desc *desc,
                                     const void *u0, void *u_final,
                                     long long nsteps, const void *dt,
                                     int every, void *workspace,
                                     size_t workspace_bytes, void *stream);
int psa_ssprk33_adjoint(const psa_solver_desc *desc, long long nsteps,
                        const void *dt, int every, const void *lam_final,
                        void *grad_u0, void *workspace, size_t workspace_bytes,
                        void *stream);

/* The same for a batch of `batch` instances of the descriptor's grid, each
 * with its own initial state, step size and final cotangent: instance b's
 * arrays start at element b * ld of u0, u_final, lam_final and grad_u0
 * (ld >= m, ld * element size a multiple of 16 bytes), dt is a device array
 * of `batch` step sizes. Every step is one launch for the whole batch (the
 * whole-step kernel forward, the fused reverse step backwards), so many
 * gradients share each launch's latency; PSA_ERR_NOT_IMPLEMENTED where the
 * whole-step kernel does not apply. batch = 1 with ld = m is the single
 * form above (timestepping.py:155-215 applied to many instances). */
int psa_adjoint_workspace_bytes_batch(const psa_solver_desc *desc, long long batch,
                                      long long ld, long long nsteps, int every,
                                      size_t *bytes);
int psa_ssprk33_forward_checkpointed_batch(const psa_solver_desc *desc, long long batch,
                                           long long ld, const void *u0, void *u_final,
                                           long long nsteps, const void *dt, int every,
                                           void *workspace, size_t workspace_bytes,
                                           void *stream);
int psa_ssprk33_adjoint_batch(const psa_solver_desc *desc, long long batch, long long ld,
                              long long nsteps, const void *dt, int every,
                              const void *lam_final, void *grad_u0, void *workspace,
                              size_t workspace_bytes, void *stream);

/* Learned WENO weights ----------------------------------------------------
 * No counterpart in the reference, which has only the fixed WENO-JS weights
 * (weno.py:247-256). A 5 -> 32 -> 32 -> 3 perceptron (ReLU, softmax over
 * logits + ln d) maps the normalised 5-cell stencil to the three weights;
 * BF16 matrix-core products with FP32 accumulation; parameters are
 * PSA_MLP_NPARAM FP32 values packed as W1[32][5], b1[32], W2[32][32],
 * b2[32], W3[3][32], b3[3]. Over the cell's smoothness indicators, cells
 * with max beta > tau (min beta + eps) keep the WENO-JS weights, cells with
 * max beta <= sigma min beta take the ideal weights d without evaluating the
 * network, and the others the network's weights. The solvers take the
 * weights through psa_solver_desc.mlp / psa_batch_desc.mlp. */
#define PSA_MLP_NPARAM 1347
#define PSA_MLP_RECORD 12 /* floats per training sample: x[5], e[3], wjs[3], c */

/* omega (n x 3) and optionally the logits (n x 3) of n input rows (n x 5). */
int psa_mlp_weights(const float *params, const float *features, long long n,
                    float *omega, float *logits, void *stream);

/* Mean training loss over n sample records and its gradient: grad receives
 * PSA_MLP_NPARAM gradient entries followed by the loss (PSA_MLP_NPARAM + 1
 * floats). Per sample, with r = sum_k omega_k e_k:
 *   c (r^2 / (mean_k e_k^2 + delta) + lambda_js sum_k (omega_k - wjs_k)^2). */
int psa_mlp_grad_workspace_bytes(long long n, size_t *bytes);
int psa_mlp_grad(const float *params, const float *samples, long long n,
                 double lambda_js, double delta, float *grad, void *work,
                 size_t work_bytes, void *stream);

/* One Adam update (step counts from 1) of the PSA_MLP_NPARAM parameters. */
int psa_mlp_adam(float *params, const float *grad, float *m, float *v,
                 long long step, double lr, double beta1, double beta2,
                 double eps, void *stream);

/* Learned counterpart of psa_weno_reconstruct (same zero padding outside
 * [0, m)): ul, ur with the learned weights, WENO-JS in flagged cells. */
int psa_learned_reconstruct(int dtype, const float *params, double tau,
                            double sigma, const void *u, void *ul, void *ur,
                            int64_t m, double eps, void *stream);

/* Solve without allocating --------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): psa_control_init followed by psa_ssprk33_steps, with the
 * control block and both stage buffers taken from a static device array of
 * PSA_STATIC_SCRATCH_BYTES that every GPU the library runs on holds until the
 * process ends. The caller passes only the input u_in and the output u_out
 * (m = n + 2 g cells each, distinct or equal) and, optionally, result
 * (device memory, one psa_static_result: t, the dt of the last step taken,
 * the steps taken and the status after the last enqueued step; NULL to
 * skip). Arguments as psa_control_init and psa_ssprk33_steps; the results
 * are the same bits as theirs.
 * Limits: m <= PSA_STATIC_MAX_CELLS (either dtype), otherwise
 * PSA_ERR_WORKSPACE. The first call on a device resolves that device's copy
 * of the array (hipGetSymbolAddress), so make it outside stream capture.
 * Every call on one GPU shares that GPU's array and it keeps the last call's
 * contents: calls that overlap on one GPU (other streams, other host
 * threads) race. Nothing here calls hipMalloc or hipFree. */
#define PSA_STATIC_MAX_CELLS (1 << 20)
#define PSA_STATIC_SCRATCH_BYTES \
  (30976 + 2 * (size_t)PSA_STATIC_MAX_CELLS * 8) /* control + 2 stages */
typedef struct psa_static_result {
  double t;      /* time of the final state                     */
  double dt;     /* dt of the last step taken                   */
  long long m;   /* steps taken                                 */
  int status;    /* PSA_OK or PSA_ERR_NONFINITE_DT              */
  int pad_;
} psa_static_result;
int psa_ssprk33_solve_static(const psa_solver_desc *desc, const void *u_in,
                             void *u_out, double tstart, double tfinal,
                             long long maxit, int dt_mode, double theta,
                             double cfl_num, double fixed_dt, long long nsteps,
                             void *result, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PYSHOCKS_AMD_H_ */
