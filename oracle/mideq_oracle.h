/*
 * mideq_oracle.h -- CPU restatement of the ensemble ODE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product (libmideq.so and the mideq Python package)
 * never calls it; the product fails loudly when its HIP library is missing.
 *
 * What it restates (reference: DifferentialEquations.jl 8.0.3, which
 * re-exports OrdinaryDiffEq 7 / SciMLBase 3 -- Project.toml:7,10,13,18;
 * src/DifferentialEquations.jl:8-9):
 *   - Tsit5 perform_step, FSAL, embedded error, free 4th-order interpolant
 *     (Tsit5 re-exported per test/qa/qa.jl:119; used test/core.jl:14,32)
 *   - Rodas5 perform_step with analytic Jacobian, W = J - I/(gamma dt),
 *     partially pivoted LU (Rosenbrock family; Rodas5P at qa.jl:97)
 *   - ode_determine_initdt (behind auto_dt_reset!, qa.jl:135)
 *   - PI step controller, tstop clamping and snapping at tf
 *   - savevalues! for a vector saveat (behaviour pinned by core.jl:90-96),
 *     with the Tsit5 interpolant or a Rosenbrock dense output
 *   - AutoTsit5(Rodas5()) stiffness detection and switching (qa.jl:15)
 *   - EnsembleThreads fan-out over trajectories (qa.jl:56)
 * The upstream packages are not vendored in the reference; their published
 * algorithms are restated from the method papers (Tsitouras 2011, Di Marzo
 * 1993 / Hairer-Wanner II) and from OrdinaryDiffEq's documented behaviour.
 * Every function's parity anchor is listed in tests/ (order conditions,
 * closed forms, invariants, independent solvers).  Functions marked
 * "parity unpinned" have no check that could fail on a plausible mistake.
 *
 * Arithmetic convention shared with the HIP kernels (so that FP64 runs are
 * bit-identical): build with -ffp-contract=off, and every tableau sum is
 * written with explicit fma() in the order OrdinaryDiffEq's @muladd produces
 * (last product innermost, first product outermost, then fma(dt, acc, uprev)).
 */
#ifndef MIDEQ_ORACLE_H
#define MIDEQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_NMAX 128

/* model ids (same numbering as include/mideq.h) */
enum {
  ORACLE_MODEL_EXP = 0,         /* u' = p0 u                (test/core.jl:10-14)  */
  ORACLE_MODEL_LORENZ = 1,      /* sigma, rho, beta          (test/core.jl:22-27)  */
  ORACLE_MODEL_ROBERTSON = 2,   /* k1, k2, k3                (test/core.jl:39-45)  */
  ORACLE_MODEL_VDP = 3,         /* classic x'=y, y'=mu(1-x^2)y-x                   */
  ORACLE_MODEL_BRUSSELATOR = 4, /* 1-D periodic, p = (A, B, alpha), n = 2*Ngrid    */
  ORACLE_MODEL_LINEAR = 5,      /* u' = M u, M row-major in p (np = n*n)           */
  ORACLE_MODEL_PROTHERO = 6,    /* y' = lambda (y - sin t) + cos t (non-autonomous)*/
  ORACLE_MODEL_USER = 7         /* callbacks set by oracle_set_user_model            */
};

/* A run-time compiled RHS on the CPU: the tests compile the same statements
 * the device compiles (f for double and float, J row-major n*n, df/dt) and
 * hand the functions over.  The time argument is the stage time (float for
 * the FP32 form); tdep != 0 marks a non-autonomous model (Rosenbrock steps
 * then add the d_i dt df/dt terms).  One user model at a time (process-wide). */
typedef void (*oracle_user_f64)(double *du, const double *u, const double *p, double t);
typedef void (*oracle_user_f32)(float *du, const float *u, const float *p, float t);
typedef void (*oracle_user_jac)(double *J, const double *u, const double *p, double t);
typedef void (*oracle_user_tgrad)(double *dT, const double *u, const double *p, double t);
int oracle_set_user_model(int32_t n, int32_t np, oracle_user_f64 f64, oracle_user_f32 f32,
                          oracle_user_jac jac, oracle_user_tgrad tgrad, int32_t tdep);
/* Callbacks of the user model (ncb = 0 clears them): per callback the kind
 * (0 continuous, 1 discrete), direction (0 both, 1 up, -1 down), terminate
 * flag and interpolation points (0: the step end only); cond(k, u, p, t) and
 * affect(k, u, p, t) as the device compiles them. */
typedef double (*oracle_user_cond)(int32_t k, const double *u, const double *p, double t);
typedef void (*oracle_user_affect)(int32_t k, double *u, const double *p, double t);
int oracle_set_user_callbacks(int32_t ncb, const int32_t *kind, const int32_t *direction,
                              const int32_t *terminate, const int32_t *interp_points,
                              oracle_user_cond cond, oracle_user_affect affect);
/* save_positions of those callbacks (bit 0 before the affect, bit 1 after;
 * oracle_set_user_callbacks resets them to 0) */
int oracle_set_user_callback_saves(int32_t ncb, const int32_t *save_positions);
/* save_everystep for the user model (0 or 1; reset by oracle_set_user_model):
 * every accepted step's end point is saved into ev_t / ev_u beside the
 * callbacks' points.  Returns 0, or -1 for another value. */
int oracle_set_user_everystep(int32_t on);
/* Which callbacks have an affect (0: a continuous callback with none, which
 * saves no point after it); all have one until set.  -1 if ncb differs. */
int oracle_set_user_callback_affects(int32_t ncb, const int32_t *has_affect);

enum {
  ORACLE_ALG_TSIT5 = 0,
  ORACLE_ALG_RODAS5 = 1,
  ORACLE_ALG_AUTO = 2,     /* AutoTsit5(Rodas5()) */
  ORACLE_ALG_RODAS5P = 3,
  ORACLE_ALG_AUTO5P = 4    /* AutoTsit5(Rodas5P()) */
};
enum { ORACLE_PREC_FP64 = 0, ORACLE_PREC_MIXED = 1 };

/* return codes, mapped 1:1 to SciMLBase.ReturnCode (qa.jl:213) */
enum {
  ORACLE_RC_DEFAULT = 0,
  ORACLE_RC_SUCCESS = 1,
  ORACLE_RC_MAXITERS = 2,
  ORACLE_RC_DTLESSTHANMIN = 3,
  ORACLE_RC_UNSTABLE = 4,
  ORACLE_RC_DTNAN = 5,
  ORACLE_RC_INITIALFAILURE = 6,
  ORACLE_RC_TERMINATED = 7
};

typedef struct {
  int32_t alg;            /* ORACLE_ALG_* */
  int32_t precision;      /* ORACLE_PREC_* */
  double abstol, reltol;
  double dt;              /* > 0: fixed initial dt; 0: initdt heuristic */
  double dtmin;           /* > 0: DtLessThanMin when the proposed dt drops below */
  double dtmax;           /* > 0: cap; 0: |tf - t0| */
  int64_t maxiters;       /* 0: 100000 */
  const double *saveat;   /* [ns] ascending absolute times, may be NULL */
  int32_t ns;
  int32_t qsteady_default;/* 1: method default qsteady band (Tsit5 [1,1], Rodas5 [1,6/5]) */
  int32_t controller;     /* 0: PI controller as restated; 1: fast form (EEst^2, log2 domain, n <= 3) */
  int32_t norm_lanes;     /* n > 8: RMS sum as norm_lanes strided partials + pairwise tree
                             (the order of a team of that many GPU lanes); 0 = 64 */
  const double *dt0_in;   /* [ntraj] per-trajectory initial dt instead of initdt, or NULL
                             (ignored when dt > 0) */
  const int8_t *alg_hint; /* [ntraj] initial method of AUTO/AUTO5P (0 Tsit5, 1 Rosenbrock),
                             or NULL */
  int32_t save_layout;    /* ensemble saves: 0 [ns][n][ntraj] (trajectory fastest),
                             1 [ntraj][ns][n] (component fastest); n = the layout's rows
                             (the largest state size of a ragged batch) */
  /* points saved at callback events (save_positions), or NULL: trajectory
   * i's k-th at ev_t[i*ev_cap + k], ev_u[(i*ev_cap + k)*n + c] while
   * k < ev_cap; ev_n[i] (ensembles) the number produced */
  double *ev_t;
  double *ev_u;
  int32_t *ev_n;
  int32_t ev_cap;
  int32_t reserved;
  /* tstops (OrdinaryDiffEq's kwarg), ascending, or NULL: the steps land on
   * the ones strictly inside the span as on tf */
  const double *tstops;
  int32_t ntstops;
  int32_t reserved2;
} oracle_options;

typedef struct {
  double u[ORACLE_NMAX];  /* final state */
  double t;               /* final time */
  int32_t retcode;
  int32_t naccept, nreject, nf, njac;
  double dt0;             /* first step size tried */
  double dt_first_accept; /* size of the first accepted step */
  int32_t nstiff_steps;   /* accepted steps taken by Rodas5 (AUTO / RODAS5) */
  int32_t nswitch;        /* algorithm switches (AUTO) */
  int32_t first_stiff_iter; /* iteration of the first switch to Rodas5, -1 if none */
  int32_t nevent;         /* callback affects applied (user model callbacks) */
  double t_event;         /* time of the last one, NaN if none */
  int32_t nevsave;        /* points saved at events (save_positions) */
  int32_t reserved;
} oracle_result;

/* Single trajectory.  u0[n], p[np].  out_save (may be NULL) is [ns][n]. */
int oracle_solve(int32_t model, int32_t n, const double *u0, const double *p,
                 int32_t np, double t0, double tf, const oracle_options *opt,
                 double *out_save, oracle_result *res);

/* One attempted step of a traced solve: the step's start time, dt as the
 * loop header found it (set by the previous controller decision), the size
 * attempted (after the AutoSwitch factor and the dtmax / tf clamps), the
 * method used (0 Tsit5, 1 Rosenbrock), whether it was accepted, and the
 * stiffness estimate left by the attempt (eig / eigd). */
typedef struct {
  double t, dt_hdr, dt, eig;
  int32_t mode, accepted;
} oracle_trace_rec;

/* oracle_solve recording up to `cap` attempted steps into trace; returns
 * the number of attempts made (may exceed cap: only cap are stored). */
int oracle_solve_trace(int32_t model, int32_t n, const double *u0, const double *p,
                       int32_t np, double t0, double tf, const oracle_options *opt,
                       double *out_save, oracle_result *res, oracle_trace_rec *trace,
                       int32_t cap);

/* EnsembleThreads restatement: ntraj independent solves on nthreads host
 * threads with dynamic chunking, results in index order.
 * SoA layouts: u0[c*ntraj + i], p[j*ntraj + i], out_save[(k*n + c)*ntraj + i],
 * u_final[c*ntraj + i].  t0/tf: per-trajectory arrays.  Any output pointer
 * may be NULL.  nvec (may be NULL): per-trajectory state size (ragged). */
int oracle_ensemble_solve(int32_t model, int32_t n, int32_t np, int64_t ntraj,
                          const double *u0, const double *p, const double *t0,
                          const double *tf, const int32_t *nvec,
                          const oracle_options *opt, int32_t nthreads,
                          double *out_save, double *u_final, double *t_final,
                          int8_t *retcode, int32_t *naccept, int32_t *nreject,
                          int32_t *nf, double *dt0, int32_t *nstiff, double *dt_first,
                          int32_t *nevent, double *t_event);

/* ---- hooks that expose internals to the tests ---- */
int oracle_tsit5_tableau(double *c7, double *a49, double *btilde7, double *r28);
int oracle_rodas5_tableau(double *gamma, double *a64, double *C64, double *c8,
                          double *d8);
/* which: 0 Rodas5, 1 Rodas5P (tableau; one fixed step) */
int oracle_rosenbrock_step(int32_t which, int32_t model, int32_t n, const double *u,
                           const double *p, double t, double dt, double abstol, double reltol,
                           double *u1, double *eest);
int oracle_rosenbrock_tableau(int32_t which, double *gamma, double *a64, double *C64,
                              double *c8, double *d8);
/* Dense-output coefficients H[3][7] (row-major) of Rodas5 (0) / Rodas5P (1). */
int oracle_rosenbrock_dense(int32_t which, double *H21);
/* One FP64 Rosenbrock step from (t, u) with step dt, then the dense output
 * at ntheta values theta[k] into out[k*n + c]; u1 gets the step's end. */
int oracle_rosenbrock_interp(int32_t which, int32_t model, int32_t n, const double *u,
                             const double *p, double t, double dt, int32_t ntheta,
                             const double *theta, double *u1, double *out);
double oracle_ctl_log2(double x);
double oracle_ctl_exp2(double x);
double oracle_ctl_pow(double x, double y);
double oracle_fast_log2(double x);
/* 2^z for |z| <= log2(10) (the fast controller's clamped step factors) */
double oracle_fast_exp2(double z);
/* One PI-controller decision: returns 1 (accept) or 0 (reject) and the next
 * dt / qold, exactly as the integrator loop applies them. */
int oracle_controller(double EEst, double qold, double dt, double qsteady_max,
                      double *dt_next, double *qold_next);
/* LU with partial pivoting of row-major a[n*n] in place, then solve a x = b
 * in place in b.  Returns 0, or -1 when a pivot is exactly zero. */
int oracle_lu_solve(int32_t n, double *a, double *b);
void oracle_rhs(int32_t model, int32_t n, const double *u, const double *p,
                double t, double *du);
void oracle_jac(int32_t model, int32_t n, const double *u, const double *p,
                double t, double *J);
double oracle_initdt(int32_t model, int32_t n, const double *u0, const double *p,
                     double t0, double tf, const oracle_options *opt, int order);
/* One Tsit5 step (FP64) from (t, u) with step dt: writes u1, the error
 * estimate EEst and the 7 stages k[7][n].  For the known-answer tests. */
void oracle_tsit5_step(int32_t model, int32_t n, const double *u, const double *p,
                       double t, double dt, double abstol, double reltol,
                       double *u1, double *eest, double *k);
/* One Rodas5 step (FP64). */
int oracle_rodas5_step(int32_t model, int32_t n, const double *u, const double *p,
                       double t, double dt, double abstol, double reltol,
                       double *u1, double *eest);
/* Tsit5 dense output at theta in [0,1] from uprev, k[7][n], dt. */
void oracle_tsit5_interp(int32_t n, const double *uprev, const double *k,
                         double dt, double theta, double *out);

#ifdef __cplusplus
}
#endif
#endif
