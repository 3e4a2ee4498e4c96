/*
 * mideq.h -- C ABI of libmideq.so, the MI355X ensemble ODE solver.
 *
 * Drop-in boundary.  The reference path is
 *     solve(EnsembleProblem(prob; prob_func, output_func), Tsit5()|Rodas5(),
 *           EnsembleThreads(); trajectories, saveat, abstol, reltol)
 * (DifferentialEquations.jl 8.0.3 re-exports it from SciMLBase 3 /
 * OrdinaryDiffEq 7: src/DifferentialEquations.jl:8-9, Project.toml:13,18;
 * names EnsembleProblem/EnsembleThreads/solve at test/qa/qa.jl:50,56,192,
 * Tsit5 at qa.jl:119, Rodas5P at qa.jl:97, AutoTsit5 at qa.jl:15,
 * ReturnCode at qa.jl:213).  An ensemble backend plugs in by dispatching
 * SciMLBase.__solve on its EnsembleAlgorithm type; the backend's __solve
 * evaluates prob_func for every trajectory on the host, packs the problems
 * into structure-of-arrays buffers and calls mideq_ensemble_solve below,
 * then rebuilds the EnsembleSolution (output_func, rerun, reduction) on
 * the host.  This header is everything such a backend binds (ccall /
 * ctypes); no C++ or torch types cross it.
 *
 * Layouts (structure of arrays, trajectory index fastest -- Julia's
 * column-major Matrix{Float64}(undef, ntraj, n) has exactly this layout):
 *   u0[c*ntraj + i], p[j*ntraj + i], us[(k*n + c)*ntraj + i],
 *   u_final[c*ntraj + i].
 * Ownership: the caller owns every buffer; nothing is retained after return.
 * Threading: a context is used by one thread at a time; separate contexts
 * may run concurrently.  Device-entry solves of one context may be in flight
 * on several streams at once: each stream gets its own work-queue counters,
 * scratch buffers and cached graph (up to 8 streams per context; past that
 * the least recently used stream's are reused after that stream drains).
 * mideq_last_stats reports the last call's launch.  Errors: negative MIDEQ_E_* return values with a
 * message from mideq_last_error(); per-trajectory failures are not call
 * errors and go to retcode[i] (mapped 1:1 to SciMLBase.ReturnCode).
 */
#ifndef MIDEQ_H
#define MIDEQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIDEQ_VERSION 10000 /* 1.0.0 */

/* Registered right-hand sides (the shim maps a Julia f to one of these). */
typedef enum {
  MIDEQ_MODEL_EXP = 0,         /* u' = p0 u             (test/core.jl:10-14)          */
  MIDEQ_MODEL_LORENZ = 1,      /* p = (sigma, rho, beta) (test/core.jl:22-27)          */
  MIDEQ_MODEL_ROBERTSON = 2,   /* p = (k1, k2, k3)       (test/core.jl:39-45), has J   */
  MIDEQ_MODEL_VDP = 3,         /* p = (mu), x'=y, y'=mu(1-x^2)y-x, has J               */
  MIDEQ_MODEL_BRUSSELATOR = 4, /* p = (A, B, alpha), n = 2*Ngrid in [16,128] (ragged), J */
  MIDEQ_MODEL_LINEAR = 5       /* u' = M u, n = 3, p = M row-major (9), has J          */
} mideq_model;

typedef enum {
  MIDEQ_ALG_TSIT5 = 0,       /* Tsit5()                                     */
  MIDEQ_ALG_RODAS5 = 1,      /* Rodas5()                                    */
  MIDEQ_ALG_AUTO = 2,        /* AutoTsit5(Rodas5())                         */
  MIDEQ_ALG_RODAS5P = 3,     /* Rodas5P() (test/qa/qa.jl:97)                */
  MIDEQ_ALG_AUTO_RODAS5P = 4 /* AutoTsit5(Rodas5P())                        */
} mideq_alg;

/* FP64: every operation in double.  MIXED: stage sums and RHS evaluations
 * in FP32; state update, error norm, time (compensated), J, W and its LU in
 * FP64. */
typedef enum { MIDEQ_FP64 = 0, MIDEQ_MIXED = 1 } mideq_precision;

typedef enum {
  MIDEQ_OK = 0,
  MIDEQ_E_ARG = -1,
  MIDEQ_E_HIP = -2,
  MIDEQ_E_NOMEM = -3,
  MIDEQ_E_MODEL = -4,
  MIDEQ_E_UNSUPPORTED = -5,
  MIDEQ_E_NODEVICE = -6
} mideq_status;

/* Per-trajectory result, 1:1 with SciMLBase.ReturnCode. */
typedef enum {
  MIDEQ_RC_DEFAULT = 0,
  MIDEQ_RC_SUCCESS = 1,
  MIDEQ_RC_MAXITERS = 2,
  MIDEQ_RC_DTLESSTHANMIN = 3,
  MIDEQ_RC_UNSTABLE = 4,
  MIDEQ_RC_DTNAN = 5,
  MIDEQ_RC_INITIALFAILURE = 6, /* invalid per-trajectory input (e.g. ragged size out of range) */
  MIDEQ_RC_TERMINATED = 7      /* a callback terminated the trajectory (terminate!) */
} mideq_retcode;

typedef enum {
  MIDEQ_SCHED_AUTO = 0,       /* autotuned choice */
  MIDEQ_SCHED_LANE = 1,       /* one lane per trajectory, one launch  */
  MIDEQ_SCHED_PERSISTENT = 2  /* persistent lanes refilled from an atomic queue */
} mideq_schedule;

typedef struct mideq_ctx mideq_ctx;

/* One EnsembleProblem after prob_func: ntraj packed ODEProblems. */
typedef struct {
  int32_t model;        /* mideq_model */
  int32_t n;            /* states per trajectory (maximum when nvec != NULL) */
  int32_t np;           /* parameters per trajectory */
  int32_t reserved0;
  int64_t ntraj;
  const double *u0;     /* [n][ntraj] */
  const double *p;      /* [np][ntraj] */
  const double *t0;     /* [ntraj] or NULL: tspan0 for all */
  const double *tf;     /* [ntraj] or NULL: tspan1 for all.  tf < t0 integrates backward
                           (one-lane models; the Brusselator gives InitialFailure); tf == t0
                           is Success at u0; a NaN end is InitialFailure */
  double tspan0, tspan1;
  const int32_t *nvec;  /* [ntraj] ragged state sizes or NULL */
} mideq_problem;

/* solve() keyword arguments (names as in SciMLBase/OrdinaryDiffEq). */
typedef struct {
  int32_t alg;          /* mideq_alg */
  int32_t precision;    /* mideq_precision */
  double abstol, reltol;
  double dt;            /* > 0: initial dt for all; 0: initdt heuristic */
  double dtmin;         /* > 0: DtLessThanMin below it; 0: off */
  double dtmax;         /* > 0: cap; 0: |tf - t0| */
  int64_t maxiters;     /* 0: 100000 */
  const double *saveat; /* [ns] ascending absolute times or NULL (endpoint only); with
                           saveat, n * ntraj must be < 2^32 (MIDEQ_E_UNSUPPORTED) */
  int32_t ns;
  int32_t save_dtype;   /* 0: double, 1: float for `us` */
  const double *dt0;    /* [ntraj] per-trajectory initial |dt| (learned) or NULL */
  const int8_t *alg_hint; /* [ntraj] initial method for ALG_AUTO (0 Tsit5, 1 Rodas5) or NULL */
  int32_t qsteady_default; /* 1: method-default qsteady band; 0: [1,1] */
  int32_t schedule;     /* mideq_schedule */
  int32_t block;        /* threads per block, 0 = tuned */
  int32_t blocks_per_cu;/* persistent grid size per CU, 0 = tuned */
  int32_t lanes;        /* lanes per trajectory for large systems (32, 64), 0 = autotuned */
  int32_t graph;        /* 1: replay the launch sequence as a cached hipGraph */
  int32_t controller;   /* 0: PI controller as OrdinaryDiffEq states it; 1: same PI law
                           evaluated on EEst^2 in the log2 domain without divisions
                           (n <= 3 models) */
  int32_t auto_passes;  /* MIDEQ_ALG_AUTO / AUTO_RODAS5P on n <= 3 models: 0 or 1 = one
                           AutoSwitch pass; 2 = two passes (a Tsit5 pass with the
                           stiffness test solves every trajectory that never switches,
                           then the full AutoSwitch kernel solves the rest from t0).
                           Results are identical.  Two passes pay when a minority of
                           the batch switches and the schedule is lane (Van der Pol,
                           mu in [1, 1e4], 78% switching: 5.7 -> 5.0 ms); they cost
                           when all switch.  3 = trust the predictor: trajectories with
                           alg_hint 0 run plain Tsit5 (no stiffness test) until done or
                           past iter_cap[i] attempted steps; those, and the ones hinted
                           stiff, then run the full AutoSwitch kernel from t0.  A
                           trajectory finished in the first pass has Tsit5's result, which
                           equals AutoTsit5's when it never switches.  Ignored
                           elsewhere. */
  const int64_t *order; /* [ntraj] launch order or NULL (natural order).  A permutation
                           of 0..ntraj-1: entry k is the k-th trajectory scheduled, so
                           consecutive entries share a wavefront -- grouping trajectories
                           of similar cost and method cuts lane divergence.  Results stay
                           at natural positions.  The host entry rejects a non-permutation;
                           on the device entry, entries outside [0, ntraj) are skipped.
                           Overrides the built-in cost order of the Brusselator queue. */
  int32_t reduction;    /* mideq_reduction over the trajectories' output_func values */
  int32_t save_layout;  /* 0: us[(k*n + c)*ntraj + i] (trajectory fastest);
                           1: us[(i*ns + k)*n + c] (component fastest: one trajectory's saved
                           states contiguous, as Julia's per-trajectory sol.u; a point's
                           values are then one contiguous store per lane or team; needs
                           ns*n < 2^32) -- every model */
  const int32_t *iter_cap; /* [ntraj] auto_passes = 3: iteration cap of the plain Tsit5 pass
                              per trajectory (e.g. a multiple of the predicted cost; 0
                              sends it to AutoSwitch before its first step, starting in
                              Tsit5), or NULL for none */
  const double *tstops; /* [ntstops] ascending times the steps of every trajectory land on
                           (OrdinaryDiffEq's tstops; the ones strictly inside a span are
                           used), or NULL; needs a user model registered with
                           with_tstops = 1 (one-lane kernels) */
  int32_t ntstops;
  int32_t reserved;
} mideq_options;

/* Results; any pointer may be NULL. */
typedef struct {
  void *us;             /* [ns][n][ntraj] saveat values (double or float) */
  double *u_final;      /* [n][ntraj] */
  double *t_final;      /* [ntraj] */
  int8_t *retcode;      /* [ntraj] mideq_retcode */
  int32_t *naccept;     /* [ntraj] */
  int32_t *nreject;     /* [ntraj] */
  int32_t *nf;          /* [ntraj] RHS evaluations */
  int32_t *nstiff;      /* [ntraj] accepted Rodas5 steps */
  double *dt_first;     /* [ntraj] size of the first accepted step */
  int32_t *nevent;      /* [ntraj] callback affects applied (user models with callbacks) */
  double *t_event;      /* [ntraj] time of the last callback affect (NaN if none) */
  double *out;          /* [nout][ntraj] output_func values (user models with an output) */
  double *reduced;      /* [nout] the reduction of `out` over the trajectories */
  /* Points saved beside saveat, in the order of integration: at callback
   * events (mideq_callback.save_positions) and at every accepted step
   * (mideq_user_model.save_everystep); user models with callbacks or
   * save_everystep, one-lane kernels.  ev_t / ev_u may be NULL together (no
   * such points kept); slots past a trajectory's points are NaN. */
  double *ev_t;         /* [ntraj][ev_cap] times */
  double *ev_u;         /* [ntraj][ev_cap][n] states */
  int32_t *ev_n;        /* [ntraj] points produced; only the first ev_cap are stored */
  int32_t ev_cap;       /* points per trajectory the arrays hold (>= 1 with ev_t) */
  int32_t reserved;
} mideq_outputs;

/* Timing and launch record of the last solve on a context. */
typedef struct {
  double kernel_ms;     /* solver kernel time (HIP events on the solve stream) */
  double total_ms;      /* whole call including copies */
  int32_t schedule;     /* schedule actually used */
  int32_t block;
  int32_t grid;
  int32_t launches;
  int32_t lanes;        /* lanes per trajectory used */
  int32_t graph;        /* 1 if the launch was a graph replay */
} mideq_stats;

int mideq_version(void);
int mideq_ctx_create(mideq_ctx **out, int32_t device);
void mideq_ctx_destroy(mideq_ctx *ctx);
const char *mideq_last_error(const mideq_ctx *ctx);
int mideq_last_stats(const mideq_ctx *ctx, mideq_stats *out);

/* Replaces SciMLBase.__solve(::EnsembleProblem, alg, ::EnsembleThreads;
 * trajectories, saveat, abstol, reltol) for problems packed as above.
 * Host pointers; blocking.  Copies in, solves on the context's stream,
 * copies out. */
int mideq_ensemble_solve(mideq_ctx *ctx, const mideq_problem *prob,
                         const mideq_options *opt, mideq_outputs *out);

/* Same, with every pointer in prob/opt/out a device pointer (inputs already
 * resident in HBM).  Enqueued on `stream` (a hipStream_t, NULL = the
 * context's stream); returns after launch without synchronizing. */
int mideq_ensemble_solve_device(mideq_ctx *ctx, const mideq_problem *prob,
                                const mideq_options *opt, mideq_outputs *out,
                                void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mideq_ensemble_solve_device with its scratch -- the
 * work-queue counters, the two-pass AutoSwitch handoff list (8 bytes per
 * trajectory when auto_passes >= 2) and the reduction's block partials
 * (8 * nout * ceil(ntraj / 256) bytes) -- in a fixed region of
 * MIDEQ_NOALLOC_SCRATCH_BYTES inside the library, one per GPU, held until
 * the process ends.  Calls no hipMalloc or hipFree.  A solve needing more
 * scratch than the region holds is refused (MIDEQ_E_ARG) before any launch;
 * graph = 1 is refused (instantiating a graph allocates); a user model must
 * have its kernels for the configuration built by an earlier call of
 * another entry point; lanes = 0 on the Brusselator takes the context's
 * tuned choice or 64-lane teams (no tuning run).  Every call on one GPU
 * shares the region: calls that overlap in time on one GPU (any context,
 * any stream) race and must be ordered by the caller. */
#define MIDEQ_NOALLOC_SCRATCH_BYTES (8u << 20)
int mideq_ensemble_solve_device_noalloc(mideq_ctx *ctx, const mideq_problem *prob,
                                        const mideq_options *opt, mideq_outputs *out,
                                        void *stream);

/* ---- run-time compiled right-hand sides (hipRTC).  Replaces the user's
 * in-place f!(du, u, p, t) of ODEProblem(f!, u0, tspan, p) (test/core.jl:22-27,
 * test/qa/qa.jl:86) and ODEFunction(f!; jac) (qa.jl:84) for the lane kernels:
 * the shim passes f's body as C++ statements instead of a registered id. */
#define MIDEQ_USER_MODEL_BASE 1000 /* first id of a registered user model */
#define MIDEQ_USER_NMAX 16         /* states (one lane per trajectory)     */
#define MIDEQ_USER_NPMAX 16        /* parameters                           */
/* On-device reduction of the output_func values over the trajectories
 * (replaces reduction(u, batch, I) of EnsembleProblem, test/qa/qa.jl:50, for
 * the common folds).  Order: blocks of 256 consecutive trajectories folded
 * by a halving tree (element i with i + w, w = 128, 64, ..., 1; past the
 * end the fold's identity), then the block results in block order; MEAN
 * divides the sum by ntraj.  MIN / MAX ignore NaN (fmin / fmax). */
typedef enum {
  MIDEQ_REDUCE_NONE = 0, MIDEQ_REDUCE_SUM = 1, MIDEQ_REDUCE_MEAN = 2,
  MIDEQ_REDUCE_MIN = 3, MIDEQ_REDUCE_MAX = 4
} mideq_reduction;

/* Callbacks of a user model.  Replaces ContinuousCallback(condition, affect!)
 * and DiscreteCallback(condition, affect!) passed as solve(...; callback)
 * (test/qa/qa.jl:26,36; test/core.jl:60-79).  A continuous callback fires
 * where its condition changes sign inside an accepted step: the step's dense
 * output is checked at interp_points sub-intervals, the crossing located by
 * Illinois bracketing, the integrator moved back to it, and the affect
 * applied there (the earliest event of all continuous callbacks).  A discrete
 * callback fires at a step end where its condition is nonzero.  Every affect
 * re-evaluates the FSAL stage.  save_positions = (before, after), upstream's
 * default (true, true): each applied affect saves (t, u) just before and/or
 * just after it into mideq_outputs.ev_t / ev_u, beside the saveat grid (the
 * points a solution lists at its events). */
#define MIDEQ_USER_NCBMAX 8
#define MIDEQ_USER_NOUTMAX 16
typedef enum { MIDEQ_CB_CONTINUOUS = 0, MIDEQ_CB_DISCRETE = 1 } mideq_callback_kind;
typedef struct {
  int32_t kind;          /* mideq_callback_kind */
  int32_t direction;     /* continuous: 0 both crossings, 1 upcrossings (affect!),
                            -1 downcrossings (affect_neg!) */
  int32_t terminate;     /* 1: terminate! after the affect (retcode MIDEQ_RC_TERMINATED) */
  int32_t interp_points; /* continuous: sign checks per step (upstream's default is 10);
                            0 checks only the step end, as upstream.  Each check is one
                            dense-output evaluation: 10 make a Lorenz Tsit5 step ~1.7x
                            as expensive */
  const char *condition; /* double expression of u[], p[], t */
  const char *affect;    /* statements changing u[] (double) from u, p, t; NULL: none (a
                            continuous callback without one that does not terminate
                            saves no point after it: u is unmodified, as upstream's
                            affect! = nothing) */
  int32_t save_positions; /* bit 0: save (t, u) before the affect, bit 1: after it
                             (upstream's (true, true) is 3); needs mideq_outputs.ev_t */
  int32_t reserved;
} mideq_callback;

typedef struct {
  int32_t n;        /* states, 1..MIDEQ_USER_NMAX */
  int32_t np;       /* parameters, 0..MIDEQ_USER_NPMAX */
  const char *rhs;  /* statements of u' = f(u, p, t): assign du[0..n-1] from u[i], p[j]
                       and the time t (a model whose statements name t is non-autonomous:
                       t is the stage time, and Rosenbrock steps add d_i dt df/dt with
                       df/dt by forward-mode differentiation).  The number type is T
                       (double; float for FP32 stages; a dual number when
                       differentiated), t included: write T(2) or 2.0 for constants and
                       use sqrt exp log sin cos tanh fabs pow(x, c). */
  const char *jac;  /* statements assigning double J[i][j] = df_i/du_j from u[], p[], t;
                       NULL: forward-mode automatic differentiation of rhs (the default
                       of OrdinaryDiffEq's Rosenbrock methods without a Jacobian) */
  int32_t ncallbacks;                /* 0..MIDEQ_USER_NCBMAX */
  int32_t save_everystep;            /* 1: save_everystep = true -- every accepted step's end
                                        (t, u) is saved into mideq_outputs.ev_t / ev_u with
                                        the callbacks' points, as OrdinaryDiffEq's
                                        savevalues! pushes it: not when the last saved point
                                        (saveat or extra) has the same time, and in a step
                                        where a callback is applied, once, before the first
                                        affect (its save_positions[1] point then is not
                                        repeated).  (t0, u0) is the solution's first point.
                                        0: off */
  const mideq_callback *callbacks;   /* [ncallbacks] or NULL */
  int32_t nout;       /* on-device output_func: values per trajectory, 0..MIDEQ_USER_NOUTMAX */
  int32_t with_tstops; /* 1: the kernels honour mideq_options.tstops (steps land on them,
                          as on tf); 0: tstops refused */
  const char *output; /* statements assigning double out[0..nout-1] from the final state
                         u[], p[], the final time t and the retcode rc; replaces
                         output_func(sol, i) -> (out, false) of EnsembleProblem
                         (test/qa/qa.jl:50) for values of the final state */
} mideq_user_model;

/* Registers and compiles a user model on the context's device; *model_id
 * (>= MIDEQ_USER_MODEL_BASE) then goes in mideq_problem.model, with n and np
 * as registered.  A compile error returns MIDEQ_E_ARG with the compiler log
 * in mideq_last_error.  Kernels for further configurations (algorithm,
 * precision, controller, dense output) are compiled on their first use. */
int mideq_model_register(mideq_ctx *ctx, const mideq_user_model *m, int32_t *model_id);

/* Compiles a user model for gfx950 without a context or device (a syntax
 * check of f for double, float and dual numbers and of jac); `log` (may be
 * NULL) receives the compiler output.  Returns 0 or MIDEQ_E_ARG. */
int mideq_model_check(const mideq_user_model *m, char *log, size_t log_size);

/* 1 if statement text names the time t -- the rule by which
 * mideq_model_register makes a model non-autonomous -- else 0.  t is
 * reserved for the time: comments and character/string literals are not
 * read, and a model that declares its own local t shadows it (the model is
 * then treated as non-autonomous with df/dt = 0).  No counterpart in the
 * reference; bindings use it to report how a model will be compiled. */
int mideq_uses_time(const char *text);

/* ---- learned dt0 / method-class / cost predictor (no counterpart in the
 * reference: it feeds dt0, alg_hint, iter_cap and order above; the initdt
 * heuristic and AutoTsit5 switching remain the fallback).  16 -> 64 -> 64 ->
 * 64 -> 4 GELU MLP, BF16 weights/activations, FP32 accumulation on the
 * matrix cores.  Parameters (float32, W[in][out] row-major): W1[16x64]
 * b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[64x4] b4[4].  Outputs per
 * sample: (log10 dt0, stiffness logit, log2 cost, never-stiff logit).  X, Y,
 * target, label, cost, never are device pointers. */
#define MIDEQ_MLP_NIN 16
#define MIDEQ_MLP_NOUT 4
#define MIDEQ_MLP_NPARAM 9668
typedef struct mideq_mlp mideq_mlp;
int mideq_mlp_create(mideq_mlp **out, int32_t device);
void mideq_mlp_destroy(mideq_mlp *m);
const char *mideq_mlp_last_error(const mideq_mlp *m);
int mideq_mlp_set_params(mideq_mlp *m, const float *params);   /* host [NPARAM] */
int mideq_mlp_get_params(mideq_mlp *m, float *params);         /* host [NPARAM] */
int mideq_mlp_set_norm(mideq_mlp *m, const float *mean, const float *std); /* host [16] */
/* Y[B][4] = net(X[B][16]) */
int mideq_mlp_infer(mideq_mlp *m, const float *X, int64_t B, float *Y, void *stream);
/* gradient of w_dt*mean((y0-target)^2) + w_cls*mean(BCE(y1,label))
 * + w_cost*mean((y2-cost)^2) + w_never*mean(BCE(y3,never)) into the
 * internal buffer (cost / never may be NULL: that term is dropped);
 * loss_host (may be NULL) receives the loss (synchronizes) */
int mideq_mlp_grad(mideq_mlp *m, const float *X, const float *target, const float *label,
                   const float *cost, const float *never, int64_t B, float w_dt, float w_cls,
                   float w_cost, float w_never, float *loss_host, void *stream);
int mideq_mlp_get_grads(mideq_mlp *m, float *grads);           /* host [NPARAM] */
/* order[B] (device) = a launch order for mideq_options.order from Y[B][4]
 * (device, the output of mideq_mlp_infer): predicted non-stiff trajectories
 * first, each class by decreasing predicted cost in 1/64-octave buckets
 * (order inside a bucket arbitrary).  Enqueued on `stream`. */
int mideq_mlp_launch_order(mideq_mlp *m, const float *Y, int64_t B, int64_t *order,
                           void *stream);
int mideq_mlp_adam_step(mideq_mlp *m, float lr, float beta1, float beta2, float eps,
                        void *stream);

#ifdef __cplusplus
}
#endif
#endif
