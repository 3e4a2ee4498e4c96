/*
 * mideq_oracle.c -- CPU restatement of the ensemble ODE hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see mideq_oracle.h): the checker for the HIP
 * kernels and the EnsembleThreads CPU baseline, never part of the product.
 *
 * Build: oracle/Makefile (gcc -O2 -mfma -ffp-contract=off).  Explicit fma()
 * marks every place where OrdinaryDiffEq's @muladd contracts a product into
 * a sum; everything else is separately rounded.
 *
 * Parity status per function (anchors in tests/test_oracle_*.py):
 *   tsit5 tableau / step / interpolant ... pinned: RK order conditions to
 *        order 5 (embedded 4), interpolant conditions, closed form of
 *        u'=1.01u (test/core.jl:10-14), expm of linear systems, DOP853.
 *   rodas5 tableau / step ................ pinned: Rosenbrock order
 *        conditions to order 5 (embedded 4), c_i/d_i row sums, stiff
 *        accuracy, R(inf)=0, Prothero-Robinson and Robertson (sum y = 1,
 *        Radau cross-check).
 *   LU ................................... pinned: numpy.linalg.solve.
 *   ctl_log2/exp2/pow .................... pinned: math.log2 / 2**x bounds.
 *   initdt ............................... pinned: hand-derived d0,d1,d2
 *        for u'=1.01u (HNW II.4 algorithm).
 *   Rosenbrock dense output (saveat) ..... pinned: its defining conditions
 *        recomputed from the tableau (order 4 for every theta, stiff
 *        limit), measured local order 4 vs DOP853, exact end points; its
 *        values vs upstream's own Rodas interpolant: parity unpinned (the
 *        upstream H table is not vendored; this is the library's own
 *        derivation).
 *   PI controller, tstop snap, saveat .... pinned: hand-computed single
 *        steps, 11-point saveat count and exact u0 copy (core.jl:34,95).
 *   AutoSwitch rule ...................... pinned: on u' = lambda u and
 *        on a lambda(t) whose stiffness drops, the traced method / dt
 *        sequence equals a replay of the rule (stability size 3.5068, 9/10,
 *        10 / 3 consecutive tests, dtfac 2) with the exact estimate
 *        |lambda|; the estimates themselves equal |lambda| (Tsit5 to
 *        rounding at the step end, |J|_inf exactly at the step start).
 *        switch_max: not restated (see the AutoSwitch constants below).
 *   accept/reject sequences vs upstream .. parity unpinned (the upstream
 *        FastPower approximation and Julia's muladd association cannot be
 *        observed without Julia; the restatement fixes its own).
 */
#include "mideq_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define NMAX ORACLE_NMAX

/* ------------------------------------------------------------------------ */
/* Tsit5 tableau: Tsitouras, Comput. Math. Appl. 62 (2011) 770-775,          */
/* as tabulated by OrdinaryDiffEq (Tsit5ConstantCacheActual).                */
/* ------------------------------------------------------------------------ */
static const double TS_C[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097,
                               1.0, 1.0};
static const double TS_A[7][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365,
     -0.09249506636175525, 0, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159,
     -0.071584973281401, -0.028269050394068383, 0, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
     -3.290069515436081, 2.324710524099774, 0}};
/* btilde = bhat - b (embedded 4th order minus 5th order weights) */
static const double TS_BT[7] = {-0.00178001105222577714, -0.0008164344596567469,
                                0.007880878010261995,    -0.1447110071732629,
                                0.5823571654525552,      -0.45808210592918697,
                                0.015151515151515152};
/* dense output b_i(theta) = r_i1 theta + r_i2 theta^2 + r_i3 theta^3 + r_i4 theta^4 */
static const double TS_R[7][4] = {
    {1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
    {0.0, 0.13169999999999998, -0.2234, 0.1017},
    {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
    {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
    {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
    {0.0, 1.5, -4.0, 2.5}};

/* ------------------------------------------------------------------------ */
/* Rodas5 tableau (Di Marzo 1993; Hairer-Wanner II), transformed W-form as   */
/* tabulated by OrdinaryDiffEq (Rodas5Tableau): a = alpha Gamma^-1,          */
/* C = diag(1/gamma) - Gamma^-1.  Stages 7 and 8 reuse row 6 of a plus       */
/* unit entries (stiffly accurate), so only rows 2..6 are stored.            */
/* ------------------------------------------------------------------------ */
typedef struct {
  double gamma;
  double A[6][5];
  double C[8][7];
  double c[8];
  double d[8];
  /* continuous extension (this library's derivation, not an upstream table):
   * u(th) = (1-th) uprev + th (u1 + (1-th)(K1 + th (K2 + th K3))),
   * K_m = sum_{j<7} H[m][j] k_j.  For each m, H solves the 8 Rosenbrock
   * order conditions up to order 4 for every th (rank 6), the stiff limit
   * u(th)/uprev -> (1-th)^4 for y' = lambda y, lambda dt -> -inf, and puts
   * no weight on the error stage k_8: a consistent 9x7 system of rank 7. */
  double H[3][7];
} rdtab_t;

static const rdtab_t RD5 = {
    0.19,
    {{0, 0, 0, 0, 0},
     {2.0, 0, 0, 0, 0},
     {3.040894194418781, 1.041747909077569, 0, 0, 0},
     {2.576417536461461, 1.622083060776640, -0.9089668560264532, 0, 0},
     {2.760842080225597, 1.446624659844071, -0.3036980084553738, 0.2877498600325443, 0},
     {-14.09640773051259, 6.925207756232704, -41.47510893210728, 2.343771018586405,
      24.13215229196062}},
    {{0, 0, 0, 0, 0, 0, 0},
     {-10.31323885133993, 0, 0, 0, 0, 0, 0},
     {-21.04823117650003, -7.234992135176716, 0, 0, 0, 0, 0},
     {32.22751541853323, -4.943732386540191, 19.44922031041879, 0, 0, 0, 0},
     {-20.69865579590063, -8.816374604402768, 1.260436877740897, -0.7495647613787146, 0, 0, 0},
     {-46.22004352711257, -17.49534862857472, -289.6389582892057, 93.60855400400906,
      318.3822534212147, 0, 0},
     {34.20013733472935, -14.15535402717690, 57.82335640988400, 25.83362985412365,
      1.408950972071624, -6.551835421242162, 0},
     {42.57076742291101, -13.80770672017997, 93.98938432427124, 18.77919633714503,
      -31.58359187223370, -6.685968952921985, -5.810979938412932}},
    {0.0, 0.38, 0.3878509998321533, 0.4839718937873840, 0.4570477008819580, 1.0, 1.0, 1.0},
    {0.19, -0.1823079225333714636, -0.319231832186874912, 0.3449828624712811,
     -0.377417564392089818, 0.0, 0.0, 0.0},
    {{46.335061333323615, -6.136914790748748, 108.40712618588684, -15.432459674306823,
      -79.5099979954734, -1.9050165705652515, -2.7727981784318065},
     {-39.14106358101113, -3.4608988577808986, -157.10746284880244, 34.66137042565768,
      146.53644774500066, 4.746961424327032, 8.03303264503124},
     {35.36683737689737, 3.300351318514294, 162.0764418691076, -9.977463628946106,
      -129.9992347337098, -8.753639789000294, -10.993114152187328}}};

/* Rodas5P (Steinebach, "Construction of Rosenbrock-Wanner method Rodas5P
 * and numerical benchmarks within the Julia Differential Equations
 * package", BIT 63 (2023) 27), same W-form and stage structure as Rodas5:
 * the Rosenbrock that DifferentialEquations re-exports (test/qa/qa.jl:97,
 * src/DifferentialEquations.jl:28). */
static const rdtab_t RD5P = {
    0.21193756319429014,
    {{0, 0, 0, 0, 0},
     {3.0, 0, 0, 0, 0},
     {2.849394379747939, 0.45842242204463923, 0, 0, 0},
     {-6.954028509809101, 2.489845061869568, -10.358996098473584, 0, 0},
     {2.8029986275628964, 0.5072464736228206, -0.3988312541770524, -0.04721187230404641, 0},
     {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
      0.030198172008377946}},
    {{0, 0, 0, 0, 0, 0, 0},
     {-14.155112264123755, 0, 0, 0, 0, 0, 0},
     {-17.97296035885952, -2.859693295451294, 0, 0, 0, 0, 0},
     {147.12150275711716, -1.41221402718213, 71.68940251302358, 0, 0, 0, 0},
     {165.43517024871676, -0.4592823456491126, 42.90938336958603, -5.961986721573306, 0, 0, 0},
     {24.854864614690072, -3.0009227002832186, 47.4931110020768, 5.5814197821558125,
      -0.6610691825249471, 0, 0},
     {30.91273214028599, -3.1208243349937974, 77.79954646070892, 34.28646028294783,
      -19.097331116725623, -28.087943162872662, 0},
     {37.80277123390563, -3.2571969029072276, 112.26918849496327, 66.9347231244047,
      -40.06618937091002, -54.66780262877968, -9.48861652309627}},
    {0.0, 0.6358126895828704, 0.4095798393397535, 0.9769306725060716, 0.4288403609558664, 1.0,
     1.0, 1.0},
    {0.21193756319429014, -0.42387512638858027, -0.3384627126235924, 1.8046452872882734,
     2.325825639765069, 0.0, 0.0, 0.0},
    {{18.16834780539935, -2.4039764856508183, -28.49033949791586, -39.2353239248467,
      24.203612973337464, 31.138884328789313, 5.813992390809251},
     {20.204639179440537, -1.56515746765803, 153.73045429293606, 118.22912892648397,
      -71.49081351328653, -100.21520495678435, -27.642530934779916},
     {-15.086356891234155, 3.4125953268012186, -59.68473410632594, -45.47407592791441,
      28.595178703267447, 42.46386356049424, 18.195502534440195}}};

/* ------------------------------------------------------------------------ */
/* Controller constants (OrdinaryDiffEq defaults for an order-5 method).     */
/* ------------------------------------------------------------------------ */
#define CTL_BETA1 (7.0 / 50.0)  /* 7/(10*order) */
#define CTL_BETA2 (2.0 / 25.0)  /* 2/(5*order)  */
#define CTL_GAMMA 0.9
#define CTL_QMIN 0.2
#define CTL_QMAX 10.0
#define CTL_QOLDINIT 1e-4
#define AS_STAB_SIZE 3.5068     /* Tsit5 stability size on the real axis */
#define AS_MAXSTIFF 10
#define AS_MAXNONSTIFF 3
#define AS_TOL 0.9
#define AS_TOL_S (AS_TOL * AS_STAB_SIZE) /* fast form of the stiffness test */
#define AS_DTFAC 2.0
/* Upstream's AutoSwitch also takes switch_max = 5.  Its role (a bound on
 * consecutive non-stiff test results that decides whether a step's error
 * test may be skipped) cannot be checked against the source here; this
 * restatement runs every step's error test and never caps the number of
 * switches.  Counted per trajectory in oracle_result.nswitch. */

/* ------------------------------------------------------------------------ */
/* Controller power: x^y = exp2(y log2 x) with own series (atanh series for  */
/* ln of the mantissa, Taylor series for 2^f on |f| <= 1/2).  ~1e-11 rel.    */
/* Only +, *, fma, one division, frexp, ldexp, rint: identical on the GPU.   */
/* ------------------------------------------------------------------------ */
#define LOG2E_D 1.4426950408889634074
#define LN2_D 0.69314718055994530942
#define SQRT_HALF_D 0.70710678118654752440
#define LOG2_10_D 3.32192809488736234787

double oracle_ctl_log2(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return INFINITY;
  int e;
  double m = frexp(x, &e);
  if (m < SQRT_HALF_D) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = fma(s2, 1.0 / 13.0, 1.0 / 11.0);
  p = fma(s2, p, 1.0 / 9.0);
  p = fma(s2, p, 1.0 / 7.0);
  p = fma(s2, p, 1.0 / 5.0);
  p = fma(s2, p, 1.0 / 3.0);
  p = fma(s2, p, 1.0);
  double lnm = (2.0 * s) * p;
  return fma(lnm, LOG2E_D, (double)e);
}

double oracle_ctl_exp2(double z) {
  if (z != z) return z;
  if (z > 1020.0) return INFINITY;
  if (z < -1020.0) return 0.0;
  double nn = rint(z);
  double f = z - nn;
  double x = f * LN2_D;
  double p = fma(x, 1.0 / 362880.0, 1.0 / 40320.0);
  p = fma(x, p, 1.0 / 5040.0);
  p = fma(x, p, 1.0 / 720.0);
  p = fma(x, p, 1.0 / 120.0);
  p = fma(x, p, 1.0 / 24.0);
  p = fma(x, p, 1.0 / 6.0);
  p = fma(x, p, 0.5);
  p = fma(x, p, 1.0);
  p = fma(x, p, 1.0);
  return ldexp(p, (int)nn);
}

double oracle_ctl_pow(double x, double y) {
  return oracle_ctl_exp2(y * oracle_ctl_log2(x));
}

/* Division-free log2 for the fast controller: log2(1 + t) = t * q(t) on
 * t = m - 1, m in [sqrt(1/2), sqrt(2)), q a degree-8 least-squares fit at
 * Chebyshev nodes (|error| < 1.4e-8). */
static const double FL2[9] = {1.442695003764732, -0.7213473333598942, 0.4809106099449815,
                              -0.36070482809673765, 0.28791832447651855, -0.23892033981765542,
                              0.2156698589589577, -0.20742103505566906, 0.12614846624505};

double oracle_fast_log2(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return INFINITY;
  int e;
  double m = frexp(x, &e);
  if (m < SQRT_HALF_D) {
    m = m * 2.0;
    e = e - 1;
  }
  double t = m - 1.0;
  double q = FL2[8];
  for (int i = 7; i >= 0; i--) q = fma(q, t, FL2[i]);
  return fma(t, q, (double)e);
}

/* 2^z for the fast controller's step factors, whose exponent is clamped to
 * |z| <= log2(10): z = n + f with n = rint(z), |f| <= 1/2, and a degree-6
 * Taylor polynomial of e^x at x = f ln 2 (|x| <= 0.347, relative error
 * < 2e-7, exact at integer z).  No range checks: the domain is bounded. */
double oracle_fast_exp2(double z) {
  double nn = rint(z);
  double f = z - nn;
  double x = f * LN2_D;
  double p = fma(x, 1.0 / 720.0, 1.0 / 120.0);
  p = fma(x, p, 1.0 / 24.0);
  p = fma(x, p, 1.0 / 6.0);
  p = fma(x, p, 0.5);
  p = fma(x, p, 1.0);
  p = fma(x, p, 1.0);
  return ldexp(p, (int)nn);
}

/* log2 of the controller constants */
#define L2_QMAXINV (-3.321928094887362)   /* log2(1/10) */
#define L2_QMININV 2.321928094887362      /* log2(5) */
#define L2_GAMMA (-0.15200309344504997)   /* log2(0.9) */
#define L2_QOLDINIT (-13.287712379549449) /* log2(1e-4) */

/* PI controller (stepsize_controller! for PIController):                    */
/*   q11 = EEst^beta1, q = clamp(q11 / qold^beta2 / gamma, 1/qmax, 1/qmin),   */
/*   EEst == 0 gives q = 1/qmax.                                              */
static double pi_stepsize(double EEst, double qold, double *q11) {
  if (EEst == 0.0) return 1.0 / CTL_QMAX;
  *q11 = oracle_ctl_pow(EEst, CTL_BETA1);
  double q = *q11 / oracle_ctl_pow(qold, CTL_BETA2);
  return fmax(1.0 / CTL_QMAX, fmin(1.0 / CTL_QMIN, q / CTL_GAMMA));
}

int oracle_controller(double EEst, double qold, double dt, double qsteady_max,
                      double *dt_next, double *qold_next) {
  double q11 = 1.0;
  double q = pi_stepsize(EEst, qold, &q11);
  if (EEst <= 1.0) {
    if (1.0 <= q && q <= qsteady_max) q = 1.0;
    *qold_next = fmax(EEst, CTL_QOLDINIT);
    *dt_next = dt / q;
    return 1;
  }
  *qold_next = qold;
  *dt_next = dt / fmin(1.0 / CTL_QMIN, q11 / CTL_GAMMA);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Models: RHS f(u,p,t) in FP64 and FP32, analytic Jacobian, tgrad.          */
/* Operation order follows the Julia source of each RHS literally           */
/* (left-to-right binary operators, x^2 as x*x, no contraction).            */
/* ------------------------------------------------------------------------ */
typedef struct {
  int model, n, np;
  double p[16];
  const double *pl; /* LINEAR: full matrix */
  float pf[16];
} mdl_t;

static struct {
  int n, np;
  oracle_user_f64 f64;
  oracle_user_f32 f32;
  oracle_user_jac jac;
  oracle_user_tgrad tgrad;
  int tdep;
  int ncb;
  int kind[8], dir[8], term[8], ip[8];
  int save[8]; /* save_positions bits: 1 before the affect, 2 after */
  int everystep; /* save_everystep: every accepted step's end point saved */
  int has_affect[8]; /* 0: a continuous callback without an affect (u unmodified) */
  oracle_user_cond cond;
  oracle_user_affect affect;
} g_user;

int oracle_set_user_callbacks(int32_t ncb, const int32_t *kind, const int32_t *direction,
                              const int32_t *terminate, const int32_t *interp_points,
                              oracle_user_cond cond, oracle_user_affect affect) {
  if (ncb < 0 || ncb > 8 || (ncb > 0 && (!cond || !affect))) return -1;
  g_user.ncb = ncb;
  for (int k = 0; k < ncb; k++) {
    g_user.kind[k] = kind[k];
    g_user.dir[k] = direction[k];
    g_user.term[k] = terminate[k];
    g_user.ip[k] = interp_points[k] > 0 ? interp_points[k] : 1; /* 0: the step end only */
  }
  g_user.cond = cond;
  g_user.affect = affect;
  for (int k = 0; k < 8; k++) g_user.save[k] = 0;
  for (int k = 0; k < 8; k++) g_user.has_affect[k] = 1;
  return 0;
}

int oracle_set_user_callback_affects(int32_t ncb, const int32_t *has_affect) {
  if (ncb != g_user.ncb) return -1;
  for (int k = 0; k < ncb; k++) g_user.has_affect[k] = has_affect[k] != 0;
  return 0;
}

/* save_positions of callback k as applied: a continuous callback without an
 * affect (and not terminating -- terminate! is then upstream's affect)
 * leaves u unmodified, and DiffEqBase's apply_callback! then returns before
 * the save_positions[2] point (discrete callbacks save it always). */
static int cb_saves(int k) {
  return (g_user.kind[k] == 0 && !g_user.has_affect[k] && !g_user.term[k])
             ? (g_user.save[k] & 1)
             : g_user.save[k];
}

int oracle_set_user_everystep(int32_t on) {
  if (on != 0 && on != 1) return -1;
  g_user.everystep = on;
  return 0;
}

int oracle_set_user_callback_saves(int32_t ncb, const int32_t *save_positions) {
  if (ncb != g_user.ncb) return -1;
  for (int k = 0; k < ncb; k++) {
    if (save_positions[k] < 0 || save_positions[k] > 3) return -1;
    g_user.save[k] = save_positions[k];
  }
  return 0;
}

/* A point saved at a callback event (save_positions): (t, u) as point k of
 * trajectory idx while k < ev_cap; the count goes on past the capacity. */
static void save_event(const oracle_options *o, int64_t idx, int n, double t, const double *u,
                       oracle_result *res) {
  if (!o->ev_t) return;
  if (res->nevsave < o->ev_cap) {
    const long long at = (long long)idx * o->ev_cap + res->nevsave;
    o->ev_t[at] = t;
    for (int c = 0; c < n; c++) o->ev_u[at * n + c] = u[c];
  }
  res->nevsave++;
}

/* savevalues! of a step end, not forced (OrdinaryDiffEq integrator_utils),
 * after the saveat points up to t are stored: save_everystep pushes (t, u)
 * unless the last saved point has time t.  Returns savedexactly -- this call
 * saved a point at t; a saveat point at t counts in the first call of the
 * step only (the queue holds no more points <= t after it). */
static int step_point(const oracle_options *o, int64_t idx, int n, double t, const double *u,
                      int everystep, int *exact, double *tlast, oracle_result *res) {
  int se = *exact;
  *exact = 0;
  if (everystep && t != *tlast) {
    save_event(o, idx, n, t, u, res);
    *tlast = t;
    se = 1;
  }
  return se;
}

int oracle_set_user_model(int32_t n, int32_t np, oracle_user_f64 f64, oracle_user_f32 f32,
                          oracle_user_jac jac, oracle_user_tgrad tgrad, int32_t tdep) {
  if (n < 1 || n > NMAX || np < 0 || np > 16 || !f64 || (tdep && !tgrad)) return -1;
  g_user.n = n;
  g_user.np = np;
  g_user.f64 = f64;
  g_user.f32 = f32;
  g_user.jac = jac;
  g_user.tgrad = tgrad;
  g_user.tdep = tdep != 0;
  g_user.everystep = 0;
  return 0;
}

static void rhs64(const mdl_t *m, const double *u, double t, double *du) {
  const double *p = m->p;
  switch (m->model) {
    case ORACLE_MODEL_USER:
      g_user.f64(du, u, p, t);
      break;
    case ORACLE_MODEL_EXP:
      du[0] = p[0] * u[0];
      break;
    case ORACLE_MODEL_LORENZ:
      du[0] = p[0] * (u[1] - u[0]);
      du[1] = u[0] * (p[1] - u[2]) - u[1];
      du[2] = u[0] * u[1] - p[2] * u[2];
      break;
    case ORACLE_MODEL_ROBERTSON: {
      double y1 = u[0], y2 = u[1], y3 = u[2];
      double k1 = p[0], k2 = p[1], k3 = p[2];
      du[0] = -k1 * y1 + k3 * y2 * y3;
      du[1] = k1 * y1 - k2 * (y2 * y2) - k3 * y2 * y3;
      du[2] = k2 * (y2 * y2);
      break;
    }
    case ORACLE_MODEL_VDP: {
      double x = u[0], y = u[1], mu = p[0];
      du[0] = y;
      du[1] = mu * (1.0 - x * x) * y - x;
      break;
    }
    case ORACLE_MODEL_BRUSSELATOR: {
      int N = m->n / 2;
      double A = p[0], B = p[1], D = p[2] * (double)(N * N);
      double B1 = B + 1.0;
      for (int j = 0; j < N; j++) {
        int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
        double uu = u[2 * j], vv = u[2 * j + 1];
        double uuv = uu * uu * vv;
        double lu = u[2 * jm] - 2.0 * uu + u[2 * jp];
        double lv = u[2 * jm + 1] - 2.0 * vv + u[2 * jp + 1];
        du[2 * j] = A + uuv - B1 * uu + D * lu;
        du[2 * j + 1] = B * uu - uuv + D * lv;
      }
      break;
    }
    case ORACLE_MODEL_LINEAR: {
      int n = m->n;
      for (int i = 0; i < n; i++) {
        double acc = m->pl[i * n] * u[0];
        for (int j = 1; j < n; j++) acc = fma(m->pl[i * n + j], u[j], acc);
        du[i] = acc;
      }
      break;
    }
    case ORACLE_MODEL_PROTHERO:
      du[0] = p[0] * (u[0] - sin(t)) + cos(t);
      break;
  }
}

static void rhs32(const mdl_t *m, const float *u, double t, float *du) {
  const float *p = m->pf;
  switch (m->model) {
    case ORACLE_MODEL_USER:
      g_user.f32(du, u, p, (float)t);
      break;
    case ORACLE_MODEL_EXP:
      du[0] = p[0] * u[0];
      break;
    case ORACLE_MODEL_LORENZ:
      du[0] = p[0] * (u[1] - u[0]);
      du[1] = u[0] * (p[1] - u[2]) - u[1];
      du[2] = u[0] * u[1] - p[2] * u[2];
      break;
    case ORACLE_MODEL_ROBERTSON: {
      float y1 = u[0], y2 = u[1], y3 = u[2];
      float k1 = p[0], k2 = p[1], k3 = p[2];
      du[0] = -k1 * y1 + k3 * y2 * y3;
      du[1] = k1 * y1 - k2 * (y2 * y2) - k3 * y2 * y3;
      du[2] = k2 * (y2 * y2);
      break;
    }
    case ORACLE_MODEL_VDP: {
      float x = u[0], y = u[1], mu = p[0];
      du[0] = y;
      du[1] = mu * (1.0f - x * x) * y - x;
      break;
    }
    case ORACLE_MODEL_BRUSSELATOR: {
      int N = m->n / 2;
      float A = p[0], B = p[1], D = p[2] * (float)(N * N);
      float B1 = B + 1.0f;
      for (int j = 0; j < N; j++) {
        int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
        float uu = u[2 * j], vv = u[2 * j + 1];
        float uuv = uu * uu * vv;
        float lu = u[2 * jm] - 2.0f * uu + u[2 * jp];
        float lv = u[2 * jm + 1] - 2.0f * vv + u[2 * jp + 1];
        du[2 * j] = A + uuv - B1 * uu + D * lu;
        du[2 * j + 1] = B * uu - uuv + D * lv;
      }
      break;
    }
    case ORACLE_MODEL_LINEAR: {
      int n = m->n;
      for (int i = 0; i < n; i++) {
        float acc = (float)m->pl[i * n] * u[0];
        for (int j = 1; j < n; j++) acc = fmaf((float)m->pl[i * n + j], u[j], acc);
        du[i] = acc;
      }
      break;
    }
    case ORACLE_MODEL_PROTHERO:
      du[0] = p[0] * (u[0] - (float)sin(t)) + (float)cos(t);
      break;
  }
}

/* J row-major n x n */
static void jac64(const mdl_t *m, const double *u, double t, double *J) {
  const double *p = m->p;
  int n = m->n;
  (void)t;
  switch (m->model) {
    case ORACLE_MODEL_USER:
      g_user.jac(J, u, p, t);
      break;
    case ORACLE_MODEL_EXP:
      J[0] = p[0];
      break;
    case ORACLE_MODEL_LORENZ:
      J[0] = -p[0]; J[1] = p[0]; J[2] = 0.0;
      J[3] = p[1] - u[2]; J[4] = -1.0; J[5] = -u[0];
      J[6] = u[1]; J[7] = u[0]; J[8] = -p[2];
      break;
    case ORACLE_MODEL_ROBERTSON: {
      double y2 = u[1], y3 = u[2], k1 = p[0], k2 = p[1], k3 = p[2];
      J[0] = -k1; J[1] = k3 * y3; J[2] = k3 * y2;
      J[3] = k1; J[4] = -2.0 * k2 * y2 - k3 * y3; J[5] = -k3 * y2;
      J[6] = 0.0; J[7] = 2.0 * k2 * y2; J[8] = 0.0;
      break;
    }
    case ORACLE_MODEL_VDP: {
      double x = u[0], y = u[1], mu = p[0];
      J[0] = 0.0; J[1] = 1.0;
      J[2] = -2.0 * mu * x * y - 1.0; J[3] = mu * (1.0 - x * x);
      break;
    }
    case ORACLE_MODEL_BRUSSELATOR: {
      int N = n / 2;
      double B = p[1], D = p[2] * (double)(N * N);
      double B1 = B + 1.0;
      memset(J, 0, sizeof(double) * (size_t)n * (size_t)n);
      for (int j = 0; j < N; j++) {
        int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
        double uu = u[2 * j], vv = u[2 * j + 1];
        int ru = 2 * j, rv = 2 * j + 1;
        double uv2 = 2.0 * uu * vv, u2 = uu * uu;
        J[ru * n + ru] += uv2 - B1 - 2.0 * D;
        J[ru * n + rv] += u2;
        J[ru * n + 2 * jm] += D;
        J[ru * n + 2 * jp] += D;
        J[rv * n + ru] += B - uv2;
        J[rv * n + rv] += -u2 - 2.0 * D;
        J[rv * n + 2 * jm + 1] += D;
        J[rv * n + 2 * jp + 1] += D;
      }
      break;
    }
    case ORACLE_MODEL_LINEAR:
      memcpy(J, m->pl, sizeof(double) * (size_t)n * (size_t)n);
      break;
    case ORACLE_MODEL_PROTHERO:
      J[0] = p[0];
      break;
  }
}

static int model_autonomous(int model) {
  if (model == ORACLE_MODEL_USER) return !g_user.tdep;
  return model != ORACLE_MODEL_PROTHERO;
}

static void tgrad64(const mdl_t *m, const double *u, double t, double *dT) {
  if (m->model == ORACLE_MODEL_PROTHERO) {
    dT[0] = -m->p[0] * cos(t) - sin(t);
  } else if (m->model == ORACLE_MODEL_USER && g_user.tdep) {
    g_user.tgrad(dT, u, m->p, t);
  } else {
    for (int i = 0; i < m->n; i++) dT[i] = 0.0;
  }
}

/* number of parameters each model reads */
static int model_np(int model, int n) {
  switch (model) {
    case ORACLE_MODEL_USER:
      return g_user.np;
    case ORACLE_MODEL_LORENZ:
    case ORACLE_MODEL_ROBERTSON:
    case ORACLE_MODEL_BRUSSELATOR:
      return 3;
    case ORACLE_MODEL_LINEAR:
      return n * n;
    default:
      return 1;
  }
}

static void mdl_init(mdl_t *m, int model, int n, const double *p, int np) {
  memset(m, 0, sizeof(*m));
  m->model = model;
  m->n = n;
  m->np = np;
  m->pl = p;
  for (int j = 0; j < np && j < 16; j++) {
    m->p[j] = p[j];
    m->pf[j] = (float)p[j];
  }
}

/* ------------------------------------------------------------------------ */
/* Norms                                                                     */
/* ------------------------------------------------------------------------ */
/* ODE_DEFAULT_NORM for arrays: sqrt(sum(abs2, x) / length(x)); summed left  */
/* to right for n <= 8 or L = 1 (a model solved in one GPU lane per          */
/* trajectory); for larger n (the interleaved (u, v) Brusselator) as         */
/* L partial sums combined by a pairwise tree (part[l] += part[l + w] for    */
/* w = L/2, ..., 1): the order in which a team of L GPU lanes reduces.       */
/* Partial l adds, in order, the components its lane owns: the (u, v) pairs  */
/* of grid points l, l + L, ... , i.e. x_{2l}, x_{2l+1}, x_{2(l+L)}, ...     */
static double sum_sq(const double *x, int n, int L) {
  double s;
  if (n <= 8 || L == 1) {
    s = x[0] * x[0];
    for (int i = 1; i < n; i++) s = s + x[i] * x[i];
  } else {
    double part[64];
    if (L <= 0 || L > 64) L = 64;
    int slots = NMAX / L; /* components per lane */
    for (int l = 0; l < L; l++) {
      double a = 0.0;
      for (int m = 0; m < slots; m++) {
        int i = 2 * (l + L * (m / 2)) + (m % 2);
        if (i < n) a = a + x[i] * x[i];
      }
      part[l] = a;
    }
    for (int w = L / 2; w >= 1; w /= 2)
      for (int l = 0; l < w; l++) part[l] = part[l] + part[l + w];
    s = part[0];
  }
  return s;
}

static double rms_norm(const double *x, int n, int L) { return sqrt(sum_sq(x, n, L) / (double)n); }

/* Lanes that reduce a norm on the GPU: user models run one trajectory per    */
/* lane (a sequential sum for every n); the Brusselator's team reduces over   */
/* the options' norm_lanes.                                                   */
static int norm_lanes(const mdl_t *m, const oracle_options *o) {
  return m->model == ORACLE_MODEL_USER ? 1 : o->norm_lanes;
}

/* Set per solve: 1 selects the fast controller, which works on EEst^2: from
 * one division over a common denominator for n <= 3, else mean(r^2) without
 * the square root.  Its division-free stiffness test applies for n <= 3. */
static __thread int g_fast_ctl = 0;

static double err_sq_fast(const double *ut, const double *u0, const double *u1, int n,
                          double abstol, double reltol) {
  double sc[3];
  for (int i = 0; i < n; i++) sc[i] = fma(fmax(fabs(u0[i]), fabs(u1[i])), reltol, abstol);
  double num, den;
  if (n == 1) {
    num = ut[0] * ut[0];
    den = sc[0] * sc[0];
  } else if (n == 2) {
    double a0 = ut[0] * sc[1], a1 = ut[1] * sc[0];
    num = a0 * a0;
    num = fma(a1, a1, num);
    double P = sc[0] * sc[1];
    den = (P * P) * 2.0;
  } else {
    double p01 = sc[0] * sc[1], p02 = sc[0] * sc[2], p12 = sc[1] * sc[2];
    double a0 = ut[0] * p12, a1 = ut[1] * p02, a2 = ut[2] * p01;
    num = a0 * a0;
    num = fma(a1, a1, num);
    num = fma(a2, a2, num);
    double P = p01 * sc[2];
    den = (P * P) * 3.0;
  }
  /* a far too large trial step can overflow the squared products: keep +inf
   * (a rejected step) instead of inf/inf; NaN stays NaN */
  if (!(num < INFINITY)) return num;
  return num / den;
}

/* calculate_residuals + internalnorm: EEst of an error vector (EEst^2 in
 * the fast-controller mode). */
static double err_norm(const double *ut, const double *u0, const double *u1,
                       int n, double abstol, double reltol, int L) {
  if (g_fast_ctl && n <= 3) return err_sq_fast(ut, u0, u1, n, abstol, reltol);
  double r[NMAX];
  for (int i = 0; i < n; i++) {
    double sc = fma(fmax(fabs(u0[i]), fabs(u1[i])), reltol, abstol);
    r[i] = ut[i] / sc;
  }
  if (g_fast_ctl) return sum_sq(r, n, L) / (double)n; /* EEst^2 for the fast controller */
  return rms_norm(r, n, L);
}

/* Julia eps(x) for finite positive x: distance to the next float. */
static double jl_eps(double x) {
  x = fabs(x);
  return nextafter(x, INFINITY) - x;
}

/* ------------------------------------------------------------------------ */
/* ode_determine_initdt (Hairer-Norsett-Wanner II.4 with upstream edits).    */
/* ------------------------------------------------------------------------ */
static double initdt_impl(const mdl_t *m, const double *u0, const double *f0,
                          double t0, double tdir, double dtmax, const oracle_options *o,
                          int order, int *nf) {
  int n = m->n;
  double dtmin = nextafter(fmax(o->dtmin, jl_eps(t0)), INFINITY);
  double smalldt = fmax(dtmin, 1e-6);
  double sk[NMAX], tmp[NMAX];
  for (int i = 0; i < n; i++) sk[i] = fma(fabs(u0[i]), o->reltol, o->abstol);
  for (int i = 0; i < n; i++) tmp[i] = u0[i] / sk[i];
  double d0 = rms_norm(tmp, n, norm_lanes(m, o));
  for (int i = 0; i < n; i++) tmp[i] = f0[i] / sk[i];
  double d1 = rms_norm(tmp, n, norm_lanes(m, o));
  if (d1 != d1) return dtmin;
  double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : (d0 / d1) / 100.0;
  dt0 = fmin(dt0, dtmax);
  if (dt0 < 10.0 * DBL_EPSILON) return smalldt;
  double u1[NMAX], f1[NMAX];
  double sdt0 = tdir * dt0; /* the trial step in the direction of integration */
  for (int i = 0; i < n; i++) u1[i] = fma(sdt0, f0[i], u0[i]);
  rhs64(m, u1, t0 + sdt0, f1);
  *nf += 1;
  int same = 1;
  for (int i = 0; i < n; i++) same &= (f0[i] == f1[i]);
  if (same) return fmax(dtmin, 100.0 * dt0);
  for (int i = 0; i < n; i++) tmp[i] = (f1[i] - f0[i]) / sk[i];
  double d2 = rms_norm(tmp, n, norm_lanes(m, o)) / dt0;
  double mx = fmax(d1, d2);
  double dt1;
  if (mx <= 1e-15) {
    dt1 = fmax(1e-6, dt0 * 1e-3);
  } else {
    /* 10^(-(2 + log10(mx))/order) */
    dt1 = oracle_ctl_exp2(-(2.0 * LOG2_10_D + oracle_ctl_log2(mx)) / (double)order);
  }
  return fmax(dtmin, fmin(fmin(100.0 * dt0, dt1), dtmax));
}

double oracle_initdt(int32_t model, int32_t n, const double *u0, const double *p,
                     double t0, double tf, const oracle_options *opt, int order) {
  mdl_t m;
  mdl_init(&m, model, n, p, model_np(model, n));
  double f0[NMAX];
  rhs64(&m, u0, t0, f0);
  int nf = 0;
  double dtmax = opt->dtmax > 0 ? opt->dtmax : fabs(tf - t0);
  return initdt_impl(&m, u0, f0, t0, tf < t0 ? -1.0 : 1.0, dtmax, opt, order, &nf);
}

/* ------------------------------------------------------------------------ */
/* Dense LU with partial pivoting (first maximal |pivot| wins), storing      */
/* multipliers below the diagonal and reciprocal pivots separately.          */
/* ------------------------------------------------------------------------ */
static int lu_factor(int n, double *a, int *piv, double *rpiv) {
  for (int k = 0; k < n; k++) {
    int pk = k;
    double best = fabs(a[k * n + k]);
    for (int i = k + 1; i < n; i++) {
      double v = fabs(a[i * n + k]);
      if (v > best) {
        best = v;
        pk = i;
      }
    }
    piv[k] = pk;
    if (pk != k) {
      for (int j = 0; j < n; j++) {
        double tt = a[k * n + j];
        a[k * n + j] = a[pk * n + j];
        a[pk * n + j] = tt;
      }
    }
    double d = a[k * n + k];
    if (d == 0.0) return -1;
    double r = 1.0 / d;
    rpiv[k] = r;
    for (int i = k + 1; i < n; i++) {
      double l = a[i * n + k] * r;
      a[i * n + k] = l;
      for (int j = k + 1; j < n; j++) a[i * n + j] = fma(-l, a[k * n + j], a[i * n + j]);
    }
  }
  return 0;
}

static void lu_solve(int n, const double *a, const int *piv, const double *rpiv,
                     double *b) {
  for (int k = 0; k < n; k++) {
    if (piv[k] != k) {
      double tt = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = tt;
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) b[i] = fma(-a[i * n + j], b[j], b[i]);
  for (int j = n - 1; j >= 0; j--) {
    b[j] = b[j] * rpiv[j];
    for (int i = 0; i < j; i++) b[i] = fma(-a[i * n + j], b[j], b[i]);
  }
}

int oracle_lu_solve(int32_t n, double *a, double *b) {
  int piv[NMAX];
  double rpiv[NMAX];
  if (lu_factor(n, a, piv, rpiv) != 0) return -1;
  lu_solve(n, a, piv, rpiv, b);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* AutoSwitch stiffness estimate max_c |dk_c / dg_c| as a fraction num/den.  */
/* Reference form: one division per component (den = 1).  Fast-controller    */
/* form: the running maximum kept as (num, den) and compared by cross-       */
/* multiplying; a component whose quotient would be NaN makes num NaN.       */
/* ------------------------------------------------------------------------ */
static void stiff_estimate(const double *dk, const double *dg, int n, double *num,
                           double *den) {
  int nan_seen = 0;
  if (!g_fast_ctl || n > 3) {
    double e = 0.0;
    for (int c = 0; c < n; c++) {
      double r = fabs(dk[c] / dg[c]);
      if (r != r) nan_seen = 1;
      if (r > e) e = r;
    }
    *num = nan_seen ? NAN : e;
    *den = 1.0;
    return;
  }
  double en = 0.0, ed = 1.0;
  for (int c = 0; c < n; c++) {
    double nn = fabs(dk[c]), dd = fabs(dg[c]);
    if (nn != nn || dd != dd || (nn == 0.0 && dd == 0.0) || (nn == INFINITY && dd == INFINITY))
      nan_seen = 1;
    if (nn * ed > en * dd) {
      en = nn;
      ed = dd;
    }
  }
  *num = nan_seen ? NAN : en;
  *den = ed;
}

/* ------------------------------------------------------------------------ */
/* Tsit5 step, FP64.  k[0] = f(uprev) on entry (FSAL); on exit k[0..6]       */
/* hold the 7 stages and u1 the new state.  Returns EEst; *eig receives the  */
/* composite-algorithm eigenvalue estimate max|(k7-k6)/(g7-g6)|.             */
/* ------------------------------------------------------------------------ */
static double tsit5_step64(const mdl_t *m, const double *uprev, double t,
                           double dt, double abstol, double reltol,
                           double k[7][NMAX], double *u1, double *eig, double *eigd, int L) {
  int n = m->n;
  double g[NMAX], g6[NMAX], ut[NMAX];
  double a = dt * TS_A[1][0];
  for (int c = 0; c < n; c++) g[c] = fma(a, k[0][c], uprev[c]);
  rhs64(m, g, t + TS_C[1] * dt, k[1]);
  for (int s = 2; s <= 6; s++) {
    for (int c = 0; c < n; c++) {
      double acc = TS_A[s][s - 1] * k[s - 1][c];
      for (int j = s - 2; j >= 0; j--) acc = fma(TS_A[s][j], k[j][c], acc);
      g[c] = fma(dt, acc, uprev[c]);
    }
    if (s == 5)
      for (int c = 0; c < n; c++) g6[c] = g[c];
    if (s < 6) {
      rhs64(m, g, t + TS_C[s] * dt, k[s]);
    } else {
      for (int c = 0; c < n; c++) u1[c] = g[c];
    }
  }
  /* stage s=5 input is g6, its RHS is k6 (index 5); s=6 row gives u1 */
  rhs64(m, u1, t + dt, k[6]);
  {
    double dk[NMAX], dg[NMAX];
    for (int c = 0; c < n; c++) {
      dk[c] = k[6][c] - k[5][c];
      dg[c] = u1[c] - g6[c];
    }
    stiff_estimate(dk, dg, n, eig, eigd);
  }
  for (int c = 0; c < n; c++) {
    double acc = TS_BT[6] * k[6][c];
    for (int j = 5; j >= 0; j--) acc = fma(TS_BT[j], k[j][c], acc);
    ut[c] = dt * acc;
  }
  return err_norm(ut, uprev, u1, n, abstol, reltol, L);
}

/* Mixed Tsit5 step: stages, stage sums and the RHS in FP32; the state      */
/* update u1 = uprev + dt*(sum b_j k_j), the error norm and t in FP64.       */
static double tsit5_step_mix(const mdl_t *m, const double *uprev, double t,
                             double dt, double abstol, double reltol,
                             float k[7][NMAX], double *u1, double *eig, double *eigd, int L) {
  int n = m->n;
  float u32[NMAX], g[NMAX], g6[NMAX], u1f[NMAX];
  double ut[NMAX];
  float dt32 = (float)dt;
  for (int c = 0; c < n; c++) u32[c] = (float)uprev[c];
  float a = dt32 * (float)TS_A[1][0];
  for (int c = 0; c < n; c++) g[c] = fmaf(a, k[0][c], u32[c]);
  rhs32(m, g, t + TS_C[1] * dt, k[1]);
  for (int s = 2; s <= 5; s++) {
    for (int c = 0; c < n; c++) {
      float acc = (float)TS_A[s][s - 1] * k[s - 1][c];
      for (int j = s - 2; j >= 0; j--) acc = fmaf((float)TS_A[s][j], k[j][c], acc);
      g[c] = fmaf(dt32, acc, u32[c]);
    }
    if (s == 5)
      for (int c = 0; c < n; c++) g6[c] = g[c];
    rhs32(m, g, t + TS_C[s] * dt, k[s]);
  }
  for (int c = 0; c < n; c++) {
    float acc = (float)TS_A[6][5] * k[5][c];
    for (int j = 4; j >= 0; j--) acc = fmaf((float)TS_A[6][j], k[j][c], acc);
    u1[c] = fma(dt, (double)acc, uprev[c]);
    u1f[c] = (float)u1[c];
  }
  rhs32(m, u1f, t + dt, k[6]);
  {
    double dk[NMAX], dg[NMAX];
    for (int c = 0; c < n; c++) {
      dk[c] = (double)k[6][c] - (double)k[5][c];
      dg[c] = u1[c] - (double)g6[c];
    }
    stiff_estimate(dk, dg, n, eig, eigd);
  }
  for (int c = 0; c < n; c++) {
    float acc = (float)TS_BT[6] * k[6][c];
    for (int j = 5; j >= 0; j--) acc = fmaf((float)TS_BT[j], k[j][c], acc);
    ut[c] = dt * (double)acc;
  }
  return err_norm(ut, uprev, u1, n, abstol, reltol, L);
}

void oracle_tsit5_step(int32_t model, int32_t n, const double *u, const double *p,
                       double t, double dt, double abstol, double reltol,
                       double *u1, double *eest, double *kout) {
  mdl_t m;
  mdl_init(&m, model, n, p, model_np(model, n));
  static __thread double k[7][NMAX];
  double eig, eigd;
  rhs64(&m, u, t, k[0]);
  *eest = tsit5_step64(&m, u, t, dt, abstol, reltol, k, u1, &eig, &eigd, 64);
  if (kout)
    for (int s = 0; s < 7; s++)
      for (int c = 0; c < n; c++) kout[s * n + c] = k[s][c];
}

/* Tsit5 free interpolant: u(theta) = uprev + dt * sum_i b_i(theta) k_i,     */
/* b_1 = theta*evalpoly(theta, r11..r14), b_i = theta^2*evalpoly(theta,      */
/* r_i2..r_i4) for i >= 2 (the upstream Tsit5 interpolant's polynomials),    */
/* evaluated in the power basis lowest power first with fma (r11 = 1):       */
/* b_1 = fma(r14, th^4, fma(r13, th^3, fma(r12, th^2, th))), b_i =           */
/* fma(ri4, th^4, fma(ri3, th^3, ri2 th^2)), th^3 = th^2 th, th^4 = th^2 th^2; */
/* the sum over stages in @muladd order.                                     */
static void tsit5_bthetas(double th, double b[7]) {
  double th2 = th * th;
  double th3 = th2 * th;
  double th4 = th2 * th2;
  b[0] = fma(TS_R[0][3], th4, fma(TS_R[0][2], th3, fma(TS_R[0][1], th2, th)));
  for (int i = 1; i < 7; i++)
    b[i] = fma(TS_R[i][3], th4, fma(TS_R[i][2], th3, TS_R[i][1] * th2));
}

static void tsit5_interp64(int n, const double *uprev, double k[7][NMAX],
                           double dt, double th, double *out) {
  double b[7];
  tsit5_bthetas(th, b);
  for (int c = 0; c < n; c++) {
    double acc = b[6] * k[6][c];
    for (int j = 5; j >= 0; j--) acc = fma(b[j], k[j][c], acc);
    out[c] = fma(dt, acc, uprev[c]);
  }
}

static void tsit5_interp_mix(int n, const double *uprev, float k[7][NMAX],
                             double dt, double th, double *out) {
  double b[7];
  float bf[7];
  tsit5_bthetas(th, b);
  for (int i = 0; i < 7; i++) bf[i] = (float)b[i];
  for (int c = 0; c < n; c++) {
    float acc = bf[6] * k[6][c];
    for (int j = 5; j >= 0; j--) acc = fmaf(bf[j], k[j][c], acc);
    out[c] = fma(dt, (double)acc, uprev[c]);
  }
}

void oracle_tsit5_interp(int32_t n, const double *uprev, const double *k,
                         double dt, double theta, double *out) {
  static __thread double kk[7][NMAX];
  for (int s = 0; s < 7; s++)
    for (int c = 0; c < n; c++) kk[s][c] = k[s * n + c];
  tsit5_interp64(n, uprev, kk, dt, theta, out);
}

/* ------------------------------------------------------------------------ */
/* Rodas5 step (FP64 or mixed).  J at uprev, W = J - I/(gamma dt) factored  */
/* once, 8 stage solves k_i = W \ (-rhs_i) with                              */
/*   U_i   = uprev + sum_{j<i} a_ij k_j        (fma ascending in j)          */
/*   rhs_i = (f(U_i) + (dt d_i) dT) + sum_{j<i} (C_ij/dt) k_j                */
/*   u     = U_6 + k6 + k7 + k8; error estimate k8.                          */
/* Returns 0 or -1 (singular W -> caller treats as EEst = 2).               */
/* ------------------------------------------------------------------------ */
typedef struct {
  double W[NMAX * NMAX];
  int piv[NMAX];
  double rpiv[NMAX];
} rodas_ws;

/* kd: the 8 stage increments (FP64; values exactly representable in FP32  */
/* when mixed), kept by the caller for the dense output.                    */
static int rodas5_step(const mdl_t *m, const rdtab_t *T, int mixed, const double *uprev,
                       double t, double dt, double abstol, double reltol, double *u1,
                       double *eest, rodas_ws *ws, double *normJ, int L,
                       double kd[8][NMAX]) {
  int n = m->n;
  double dT[NMAX], x[NMAX], U[NMAX];
  float Uf[NMAX], duf[NMAX];
  double du[NMAX];
  double dtgamma = dt * T->gamma;
  double w = 1.0 / dtgamma;
  double inv_dt = 1.0 / dt;
  int autonomous = model_autonomous(m->model);

  jac64(m, uprev, t, ws->W);
  if (normJ) {
    double best = 0.0;
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += fabs(ws->W[i * n + j]);
      if (s > best) best = s;
    }
    *normJ = best;
  }
  for (int i = 0; i < n; i++) ws->W[i * n + i] = ws->W[i * n + i] - w;
  if (lu_factor(n, ws->W, ws->piv, ws->rpiv) != 0) {
    *eest = 2.0;
    return -1;
  }
  if (!autonomous) tgrad64(m, uprev, t, dT);

  for (int s = 0; s < 8; s++) {
    /* stage input */
    double ts = t + T->c[s] * dt;
    if (!mixed) {
      if (s == 0) {
        for (int c = 0; c < n; c++) U[c] = uprev[c];
      } else if (s <= 5) {
        for (int c = 0; c < n; c++) {
          double acc = uprev[c];
          for (int j = 0; j < s; j++) acc = fma(T->A[s][j], kd[j][c], acc);
          U[c] = acc;
        }
      } else {
        /* U_7 = U_6 + k6, U_8 = U_7 + k7 */
        for (int c = 0; c < n; c++) U[c] = U[c] + kd[s - 1][c];
      }
      rhs64(m, U, ts, du);
    } else {
      if (s == 0) {
        for (int c = 0; c < n; c++) Uf[c] = (float)uprev[c];
      } else if (s <= 5) {
        for (int c = 0; c < n; c++) {
          float acc = (float)uprev[c];
          for (int j = 0; j < s; j++) acc = fmaf((float)T->A[s][j], (float)kd[j][c], acc);
          Uf[c] = acc;
        }
      } else {
        for (int c = 0; c < n; c++) Uf[c] = Uf[c] + (float)kd[s - 1][c];
      }
      rhs32(m, Uf, ts, duf);
      for (int c = 0; c < n; c++) du[c] = (double)duf[c];
    }
    /* right-hand side */
    for (int c = 0; c < n; c++) {
      double base = du[c];
      if (!autonomous) base = base + (dt * T->d[s]) * dT[c];
      double r;
      if (s == 0) {
        r = base;
      } else if (!mixed) {
        double acc = (T->C[s][0] * inv_dt) * kd[0][c];
        for (int j = 1; j < s; j++) acc = fma(T->C[s][j] * inv_dt, kd[j][c], acc);
        r = base + acc;
      } else {
        float idt = (float)inv_dt;
        float acc = ((float)T->C[s][0] * idt) * (float)kd[0][c];
        for (int j = 1; j < s; j++) acc = fmaf((float)T->C[s][j] * idt, (float)kd[j][c], acc);
        r = (double)((float)base + acc);
      }
      x[c] = -r;
    }
    lu_solve(n, ws->W, ws->piv, ws->rpiv, x);
    for (int c = 0; c < n; c++) kd[s][c] = mixed ? (double)(float)x[c] : x[c];
  }
  /* u = U_6 + k6 + k7 + k8, accumulated in FP64 in both modes */
  for (int c = 0; c < n; c++) {
    double acc = uprev[c];
    for (int j = 0; j < 5; j++) acc = fma(T->A[5][j], kd[j][c], acc);
    acc = acc + kd[5][c];
    acc = acc + kd[6][c];
    u1[c] = acc + kd[7][c];
  }
  *eest = err_norm(kd[7], uprev, u1, n, abstol, reltol, L);
  return 0;
}

/* Rosenbrock dense output at th in (0, 1) from uprev, u1 and the stage     */
/* increments (FP64 in both precisions):                                    */
/*   K_m = sum_{j<7} H[m][j] k_j   (fma ascending in j)                     */
/*   u(th) = fma(th, fma(1-th, fma(th, fma(th, K3, K2), K1), u1), (1-th) uprev) */
static void rosenbrock_interp(int n, const rdtab_t *T, const double *uprev, const double *u1,
                              double kd[8][NMAX], double th, double *out) {
  double th1 = 1.0 - th;
  for (int c = 0; c < n; c++) {
    double K[3];
    for (int mm = 0; mm < 3; mm++) {
      double acc = T->H[mm][0] * kd[0][c];
      for (int j = 1; j < 7; j++) acc = fma(T->H[mm][j], kd[j][c], acc);
      K[mm] = acc;
    }
    double inner = fma(th, K[2], K[1]);
    inner = fma(th, inner, K[0]);
    double tt = fma(th1, inner, u1[c]);
    out[c] = fma(th, tt, th1 * uprev[c]);
  }
}

int oracle_rosenbrock_step(int32_t which, int32_t model, int32_t n, const double *u,
                           const double *p, double t, double dt, double abstol, double reltol,
                           double *u1, double *eest) {
  mdl_t m;
  mdl_init(&m, model, n, p, model_np(model, n));
  rodas_ws *ws = (rodas_ws *)malloc(sizeof(rodas_ws));
  double kd[8][NMAX];
  int rc = rodas5_step(&m, which == 1 ? &RD5P : &RD5, 0, u, t, dt, abstol, reltol, u1, eest,
                       ws, NULL, 64, kd);
  free(ws);
  return rc;
}

int oracle_rosenbrock_interp(int32_t which, int32_t model, int32_t n, const double *u,
                             const double *p, double t, double dt, int32_t ntheta,
                             const double *theta, double *u1, double *out) {
  mdl_t m;
  mdl_init(&m, model, n, p, model_np(model, n));
  rodas_ws *ws = (rodas_ws *)malloc(sizeof(rodas_ws));
  double kd[8][NMAX];
  double eest;
  const rdtab_t *T = which == 1 ? &RD5P : &RD5;
  int rc = rodas5_step(&m, T, 0, u, t, dt, 1.0, 1.0, u1, &eest, ws, NULL, 64, kd);
  free(ws);
  if (rc != 0) return rc;
  for (int k = 0; k < ntheta; k++) rosenbrock_interp(n, T, u, u1, kd, theta[k], out + (long)k * n);
  return 0;
}

int oracle_rosenbrock_dense(int32_t which, double *H21) {
  const rdtab_t *T = which == 1 ? &RD5P : &RD5;
  for (int mm = 0; mm < 3; mm++)
    for (int j = 0; j < 7; j++) H21[mm * 7 + j] = T->H[mm][j];
  return 0;
}

int oracle_rodas5_step(int32_t model, int32_t n, const double *u, const double *p,
                       double t, double dt, double abstol, double reltol,
                       double *u1, double *eest) {
  return oracle_rosenbrock_step(0, model, n, u, p, t, dt, abstol, reltol, u1, eest);
}

/* Dense output of the last step at theta: Tsit5's interpolant (FP64 or    */
/* mixed) in Tsit5 mode, the Rosenbrock one in Rosenbrock mode.             */
static void step_interp(int n, int mode, int mixed, const rdtab_t *rtab, const double *uprev,
                        const double *u1, double k64[7][NMAX], float k32[7][NMAX],
                        double kd[8][NMAX], double dt, double th, double *out) {
  if (mode == 0) {
    if (!mixed)
      tsit5_interp64(n, uprev, k64, dt, th, out);
    else
      tsit5_interp_mix(n, uprev, k32, dt, th, out);
  } else {
    rosenbrock_interp(n, rtab, uprev, u1, kd, th, out);
  }
}

/* ------------------------------------------------------------------------ */
/* The integrator loop (solve! with loopheader!/perform_step!/loopfooter!). */
/* ------------------------------------------------------------------------ */
/* Trace of the attempted steps of one solve (oracle_solve_trace only). */
static __thread oracle_trace_rec *g_trace = NULL;
static __thread int32_t g_trace_cap = 0, g_trace_len = 0;

static int solve_one(const mdl_t *m, const double *u0, double t0, double tf,
                     const oracle_options *o, int64_t idx, double *save, long save_stride,
                     long save_rows, oracle_result *res, rodas_ws *ws) {
  int n = m->n;
  int mixed = o->precision == ORACLE_PREC_MIXED;
  double uprev[NMAX], u1[NMAX];
  double k64[7][NMAX];
  float k32[7][NMAX];
  double kdsave[8][NMAX]; /* Rosenbrock stage increments of the current step */
  double abstol = o->abstol, reltol = o->reltol;
  long long maxiters = o->maxiters > 0 ? o->maxiters : 100000;
  double dtmax = o->dtmax > 0 ? o->dtmax : fabs(tf - t0);
  /* 0 Tsit5, 1 Rosenbrock (Rodas5 or Rodas5P) */
  int mode = (o->alg == ORACLE_ALG_RODAS5 || o->alg == ORACLE_ALG_RODAS5P) ? 1 : 0;
  int autosw = (o->alg == ORACLE_ALG_AUTO || o->alg == ORACLE_ALG_AUTO5P);
  if (autosw && o->alg_hint) mode = o->alg_hint[idx] ? 1 : 0;
  const rdtab_t *rtab =
      (o->alg == ORACLE_ALG_RODAS5P || o->alg == ORACLE_ALG_AUTO5P) ? &RD5P : &RD5;
  double qsteady_max = (o->qsteady_default && mode == 1) ? 1.2 : 1.0;
  int count = 0;
  double eig = 0.0, eigd = 1.0; /* stiffness estimate eig / eigd */

  memset(res, 0, sizeof(*res));
  res->first_stiff_iter = -1;
  res->t_event = NAN;
  const int ncb = (m->model == ORACLE_MODEL_USER) ? g_user.ncb : 0;
  const int everystep = (m->model == ORACLE_MODEL_USER) ? g_user.everystep : 0;
  double tlast = t0; /* time of the last saved point: (t0, u0) starts the solution */
  double gprev[8];
  for (int c = 0; c < n; c++) uprev[c] = u0[c];
  double t = t0, tcomp = 0.0;
  /* direction of integration: tf < t0 integrates backward with steps of
   * -dt (dt is the step's magnitude everywhere else); the save points, in
   * ascending order, are then visited from the last one down */
  const double tdir = tf < t0 ? -1.0 : 1.0;
  const int back = tf < t0;
  /* tstops: index of the next one ahead (upstream keeps tf in the same
   * queue; tstops at or behind t0 are dropped at the start) */
  const int nts = o->tstops ? o->ntstops : 0;
  int its = back ? nts - 1 : 0;
#define TSTOP_POP                                                              \
  while (its >= 0 && its < nts && tdir * (t - o->tstops[its]) >= 0.0) its += back ? -1 : 1
#define TSTOP_NEXT                                                             \
  ((its >= 0 && its < nts && tdir * (tf - o->tstops[its]) > 0.0) ? o->tstops[its] : tf)
  const double *sa = o->saveat;
  int ns = sa ? o->ns : 0;
  int isave = back ? ns - 1 : 0;
#define SAVE_LEFT (back ? isave >= 0 : isave < ns)
#define SAVE_NEXT (back ? isave-- : isave++)
  /* save points before t0 are not reachable; the point at t0 is u0 itself */
  while (SAVE_LEFT && (back ? sa[isave] > t0 : sa[isave] < t0)) {
    if (save)
      for (int c = 0; c < n; c++) save[((long)isave * save_rows + c) * save_stride] = NAN;
    SAVE_NEXT;
  }
  while (SAVE_LEFT && sa[isave] == t0) {
    if (save)
      for (int c = 0; c < n; c++) save[((long)isave * save_rows + c) * save_stride] = u0[c];
    SAVE_NEXT;
  }

  /* FSAL initialisation: k1 = f(u0) */
  double f0[NMAX];
  rhs64(m, uprev, t, f0);
  res->nf += 1;
  if (mixed) {
    float uf[NMAX];
    for (int c = 0; c < n; c++) uf[c] = (float)uprev[c];
    rhs32(m, uf, t, k32[0]);
  }
  for (int c = 0; c < n; c++) k64[0][c] = f0[c];

  for (int k = 0; k < ncb; k++) gprev[k] = g_user.cond(k, uprev, m->p, t0);
  double dt = o->dt > 0 ? o->dt
               : o->dt0_in ? o->dt0_in[idx]
                           : initdt_impl(m, uprev, f0, t0, tdir, dtmax, o, 5, &res->nf);
  res->dt0 = dt;
  double qold = CTL_QOLDINIT;
  double q11 = 1.0;
  long long iter = 0;
  int rc = ORACLE_RC_DEFAULT;
  int have_accept = 0;
  int nanrej = 0;
  g_fast_ctl = (o->controller == 1);
  double lqold = L2_QOLDINIT;
  double lqs_cur = 0.0; /* log2 of qsteady_max: 0 or log2(1.2) */

  /* a zero-length span is done; a NaN end is not integrated */
  if (tf == t0)
    rc = ORACLE_RC_SUCCESS;
  else if (!(tf > t0) && !(tf < t0))
    rc = ORACLE_RC_INITIALFAILURE;
  while (rc == ORACLE_RC_DEFAULT && (back ? t > tf : t < tf)) {
    /* loopheader!: choose algorithm, fix dt at bounds, clamp to tstop */
    const double dt_hdr = dt;
    if (autosw && iter > 0) {
      int stiff;
      if (!g_fast_ctl || n > 3) {
        double stiffness = fabs(eig * dt / AS_STAB_SIZE);
        stiff = stiffness > AS_TOL;
      } else {
        /* |eig dt / S| > tol  <=>  num * dt > (tol S) * den */
        stiff = eig * dt > AS_TOL_S * eigd;
      }
      count = stiff ? (count < 0 ? 1 : count + 1) : (count > 0 ? -1 : count - 1);
      if (mode == 0 && count > AS_MAXSTIFF) {
        dt = dt * AS_DTFAC;
        mode = 1;
        res->nswitch++;
        if (res->first_stiff_iter < 0) res->first_stiff_iter = (int)iter;
        qsteady_max = o->qsteady_default ? 1.2 : 1.0;
      } else if (mode == 1 && count < -AS_MAXNONSTIFF) {
        dt = dt / AS_DTFAC;
        mode = 0;
        res->nswitch++;
        qsteady_max = 1.0;
        /* re-initialise FSAL for Tsit5 */
        rhs64(m, uprev, t, k64[0]);
        res->nf += 1;
        if (mixed) {
          float uf[NMAX];
          for (int c = 0; c < n; c++) uf[c] = (float)uprev[c];
          rhs32(m, uf, t, k32[0]);
        }
      }
    }
    iter++;
    if (iter > maxiters) {
      rc = ORACLE_RC_MAXITERS;
      break;
    }
    dt = fmin(dt, dtmax);
    if (o->dtmin > 0 && dt < o->dtmin) {
      rc = ORACLE_RC_DTLESSTHANMIN;
      break;
    }
    TSTOP_POP; /* handle_tstop!: the ones the last step reached */
    const double tstop = TSTOP_NEXT;
    if (nts > 0)
      dt = fmin(dt, tdir * (tstop - t)); /* modify_dt_for_tstops! */
    else
      dt = fmin(dt, back ? t - tf : tf - t);
    const double sdt = back ? -dt : dt; /* the signed step */

    /* perform_step! */
    double EEst;
    if (mode == 0) {
      if (!mixed)
        EEst = tsit5_step64(m, uprev, t, sdt, abstol, reltol, k64, u1, &eig, &eigd,
                            norm_lanes(m, o));
      else
        EEst = tsit5_step_mix(m, uprev, t, sdt, abstol, reltol, k32, u1, &eig, &eigd,
                              norm_lanes(m, o));
      res->nf += 6;
    } else {
      double nJ;
      if (rodas5_step(m, rtab, mixed, uprev, t, sdt, abstol, reltol, u1, &EEst, ws, &nJ,
                      norm_lanes(m, o), kdsave) != 0)
        EEst = 2.0;
      res->nf += 8;
      res->njac += 1;
      eig = nJ;
      eigd = 1.0;
    }
    if (EEst != EEst) {
      /* FP64: DtNaN as upstream.  Mixed: FP32 stages can overflow on a far
       * too large trial step; treat that as EEst = +inf (reject, dt/5) and
       * give up after 20 consecutive such rejections. */
      if (!mixed || ++nanrej > 20) {
        rc = ORACLE_RC_DTNAN;
        break;
      }
      EEst = INFINITY;
    } else {
      nanrej = 0;
    }

    /* stepsize_controller! (PI); the fast form works on EEst^2 in log2 */
    double q = 1.0, lq = 0.0, l11 = 0.0, lE2 = 0.0;
    lqs_cur = qsteady_max > 1.0 ? 0.2630344058337938 : 0.0;
    if (!g_fast_ctl) {
      q = pi_stepsize(EEst, qold, &q11);
    } else if (EEst == 0.0) {
      lq = L2_QMAXINV;
    } else {
      lE2 = oracle_fast_log2(EEst);
      l11 = (0.5 * CTL_BETA1) * lE2;
      lq = l11 - CTL_BETA2 * lqold - L2_GAMMA;
      lq = fmax(L2_QMAXINV, fmin(L2_QMININV, lq));
    }

    if (g_trace) {
      if (g_trace_len < g_trace_cap) {
        oracle_trace_rec *r = &g_trace[g_trace_len];
        r->t = t;
        r->dt_hdr = dt_hdr;
        r->dt = dt;
        r->eig = eig / eigd;
        r->mode = mode;
        r->accepted = EEst <= 1.0;
      }
      g_trace_len++;
    }
    if (EEst <= 1.0) {
      /* step_accept_controller! */
      res->naccept++;
      if (mode == 1) res->nstiff_steps++;
      if (!have_accept) {
        res->dt_first_accept = dt;
        have_accept = 1;
      }
      double dtnew;
      if (!g_fast_ctl) {
        if (1.0 <= q && q <= qsteady_max) q = 1.0;
        qold = fmax(EEst, CTL_QOLDINIT);
        dtnew = dt / q;
      } else {
        if (0.0 <= lq && lq <= lqs_cur) lq = 0.0;
        lqold = (EEst == 0.0) ? L2_QOLDINIT : fmax(0.5 * lE2, L2_QOLDINIT);
        dtnew = dt * oracle_fast_exp2(-lq);
      }
      double tprev = t;
      double ttmp;
      if (!mixed) {
        ttmp = t + sdt;
      } else {
        /* compensated (Kahan) time accumulation */
        double y = sdt - tcomp;
        ttmp = t + y;
        tcomp = (ttmp - t) - y;
      }
      /* fixed_t_for_floatingpoint_error!: onto the first tstop (tf included) */
      if (fabs(ttmp - tstop) < 100.0 * jl_eps(fmax(t, tstop))) ttmp = tstop;
      t = ttmp;
      /* the state the step ends in: u1, or the dense output at an event */
      double uend[NMAX];
      for (int c = 0; c < n; c++) uend[c] = u1[c];
      int ev = -1;
      if (ncb > 0) {
        /* continuous callbacks: the earliest sign change of any condition
         * over interp_points sub-intervals, Illinois (modified regula falsi)
         * bracketing on theta; the event is the bracket end past it */
        double th_ev = 2.0;
        for (int k = 0; k < ncb; k++) {
          if (g_user.kind[k] != 0) continue;
          int ip = g_user.ip[k], dir = g_user.dir[k];
          double ga = gprev[k], tha = 0.0;
          for (int q = 1; q <= ip; q++) {
            double thb, gb;
            if (q == ip) {
              thb = 1.0;
              gb = g_user.cond(k, u1, m->p, t);
            } else {
              double tmp[NMAX];
              thb = (double)q / (double)ip;
              step_interp(n, mode, mixed, rtab, uprev, u1, k64, k32, kdsave, sdt, thb, tmp);
              gb = g_user.cond(k, tmp, m->p, tprev + thb * sdt);
            }
            int up = ga < 0.0 && gb >= 0.0;
            int down = ga > 0.0 && gb <= 0.0;
            if ((up && dir >= 0) || (down && dir <= 0)) {
              double th = thb;
              if (gb != 0.0) {
                double lo = tha, hi = thb, flo = ga, fhi = gb;
                int side = 0;
                for (int it = 0; it < 100; it++) {
                  if (hi - lo <= 4.0 * 2.220446049250313e-16) break;
                  double cc = (lo * fhi - hi * flo) / (fhi - flo);
                  if (!(cc > lo && cc < hi)) cc = 0.5 * (lo + hi);
                  double tmp[NMAX];
                  step_interp(n, mode, mixed, rtab, uprev, u1, k64, k32, kdsave, sdt, cc, tmp);
                  double fc = g_user.cond(k, tmp, m->p, tprev + cc * sdt);
                  if (fc == 0.0) {
                    hi = cc;
                    break;
                  }
                  if ((fc < 0.0) == (fhi < 0.0)) {
                    hi = cc;
                    fhi = fc;
                    if (side == -1) flo = 0.5 * flo;
                    side = -1;
                  } else {
                    lo = cc;
                    flo = fc;
                    if (side == 1) fhi = 0.5 * fhi;
                    side = 1;
                  }
                }
                th = hi;
              }
              if (th < th_ev) {
                th_ev = th;
                ev = k;
              }
              break;
            }
            ga = gb;
            tha = thb;
          }
        }
        if (ev >= 0 && th_ev < 1.0) {
          t = tprev + th_ev * sdt;
          tcomp = 0.0;
          step_interp(n, mode, mixed, rtab, uprev, u1, k64, k32, kdsave, sdt, th_ev, uend);
        }
      }
      /* savevalues!: every saveat point <= t */
      int exact = 0;
      while (SAVE_LEFT && (back ? sa[isave] >= t : sa[isave] <= t)) {
        exact |= sa[isave] == t;
        tlast = sa[isave];
        if (save) {
          double val[NMAX];
          if (sa[isave] == t) {
            for (int c = 0; c < n; c++) val[c] = uend[c];
          } else {
            double th = (sa[isave] - tprev) / sdt;
            step_interp(n, mode, mixed, rtab, uprev, u1, k64, k32, kdsave, sdt, th, val);
          }
          for (int c = 0; c < n; c++) save[((long)isave * save_rows + c) * save_stride] = val[c];
        }
        SAVE_NEXT;
      }
      for (int c = 0; c < n; c++) uprev[c] = uend[c];
      if (mode == 0) {
        if (!mixed)
          for (int c = 0; c < n; c++) k64[0][c] = k64[6][c];
        else
          for (int c = 0; c < n; c++) k32[0][c] = k32[6][c];
      }
      if (ncb > 0) {
        /* apply_callback!: the continuous event's affect, then every
         * discrete callback whose condition holds; any affect re-evaluates
         * the FSAL stage */
        int fired = 0, term = 0;
        if (ev >= 0) {
          /* each applied callback: savevalues! of the step end, then the
           * point before the affect unless that call saved one at t */
          int se = step_point(o, idx, n, t, uprev, everystep, &exact, &tlast, res);
          const int sv = cb_saves(ev);
          if ((sv & 1) && !se) save_event(o, idx, n, t, uprev, res);
          g_user.affect(ev, uprev, m->p, t);
          if (sv & 2) save_event(o, idx, n, t, uprev, res);
          if (sv) tlast = t;
          fired = 1;
          term |= g_user.term[ev];
          res->nevent++;
        }
        for (int k = 0; k < ncb; k++) {
          if (g_user.kind[k] == 1 && g_user.cond(k, uprev, m->p, t) != 0.0) {
            int se = step_point(o, idx, n, t, uprev, everystep, &exact, &tlast, res);
            const int sv = cb_saves(k);
            if ((sv & 1) && !se) save_event(o, idx, n, t, uprev, res);
            g_user.affect(k, uprev, m->p, t);
            if (sv & 2) save_event(o, idx, n, t, uprev, res);
            if (sv) tlast = t;
            fired = 1;
            term |= g_user.term[k];
            res->nevent++;
          }
        }
        if (!fired) step_point(o, idx, n, t, uprev, everystep, &exact, &tlast, res);
        if (fired) {
          res->t_event = t;
          rhs64(m, uprev, t, k64[0]);
          if (mixed) {
            float uf[NMAX];
            for (int c = 0; c < n; c++) uf[c] = (float)uprev[c];
            rhs32(m, uf, t, k32[0]);
          }
          res->nf += 1;
        }
        for (int k = 0; k < ncb; k++)
          if (g_user.kind[k] == 0) gprev[k] = g_user.cond(k, uprev, m->p, t);
        if (term) {
          rc = ORACLE_RC_TERMINATED;
          break;
        }
      } else if (everystep) {
        step_point(o, idx, n, t, uprev, everystep, &exact, &tlast, res);
      }
      int bad = 0;
      for (int c = 0; c < n; c++) bad |= (u1[c] != u1[c]);
      if (bad) {
        rc = ORACLE_RC_UNSTABLE;
        break;
      }
      dt = fmin(dtmax, dtnew);
    } else {
      /* step_reject_controller! */
      res->nreject++;
      if (!g_fast_ctl)
        dt = dt / fmin(1.0 / CTL_QMIN, q11 / CTL_GAMMA);
      else
        dt = dt * oracle_fast_exp2(-fmin(L2_QMININV, l11 - L2_GAMMA));
    }
  }
  if (rc == ORACLE_RC_DEFAULT) rc = ORACLE_RC_SUCCESS;
  for (; SAVE_LEFT; SAVE_NEXT)
    if (save)
      for (int c = 0; c < n; c++) save[((long)isave * save_rows + c) * save_stride] = NAN;
#undef SAVE_LEFT
#undef SAVE_NEXT
#undef TSTOP_POP
#undef TSTOP_NEXT
  /* event-point slots past this trajectory's points hold NaN */
  if (o->ev_t)
    for (long long k = res->nevsave; k < o->ev_cap; k++) {
      const long long at = (long long)idx * o->ev_cap + k;
      o->ev_t[at] = NAN;
      for (int c = 0; c < n; c++) o->ev_u[at * n + c] = NAN;
    }
  for (int c = 0; c < n; c++) res->u[c] = uprev[c];
  res->t = t;
  res->retcode = rc;
  g_fast_ctl = 0;
  return rc;
}

int oracle_solve(int32_t model, int32_t n, const double *u0, const double *p,
                 int32_t np, double t0, double tf, const oracle_options *opt,
                 double *out_save, oracle_result *res) {
  if (n < 1 || n > NMAX) return -1;
  mdl_t m;
  mdl_init(&m, model, n, p, np);
  rodas_ws *ws = (rodas_ws *)malloc(sizeof(rodas_ws));
  if (!ws) return -2;
  solve_one(&m, u0, t0, tf, opt, 0, out_save, 1, n, res, ws);
  free(ws);
  return 0;
}

int oracle_solve_trace(int32_t model, int32_t n, const double *u0, const double *p,
                       int32_t np, double t0, double tf, const oracle_options *opt,
                       double *out_save, oracle_result *res, oracle_trace_rec *trace,
                       int32_t cap) {
  g_trace = trace;
  g_trace_cap = trace ? cap : 0;
  g_trace_len = 0;
  int rc = oracle_solve(model, n, u0, p, np, t0, tf, opt, out_save, res);
  g_trace = NULL;
  g_trace_cap = 0;
  return rc != 0 ? rc : g_trace_len;
}

/* ------------------------------------------------------------------------ */
/* EnsembleThreads restatement: a pthread pool pulling chunks of            */
/* trajectory indices from an atomic counter; outputs land in index order.  */
/* ------------------------------------------------------------------------ */
typedef struct {
  int model, n, np;
  long long ntraj;
  const double *u0, *p, *t0, *tf;
  const int32_t *nvec;
  const oracle_options *opt;
  double *out_save, *u_final, *t_final, *dt0;
  int8_t *retcode;
  int32_t *naccept, *nreject, *nf, *nstiff;
  double *dt_first;
  int32_t *nevent;
  double *t_event;
  atomic_llong next;
} ens_job;

#define ENS_CHUNK 16

static void *ens_worker(void *arg) {
  ens_job *J = (ens_job *)arg;
  rodas_ws *ws = (rodas_ws *)malloc(sizeof(rodas_ws));
  double u0[NMAX], p[16];
  double *pl = NULL;
  if (J->model == ORACLE_MODEL_LINEAR) pl = (double *)malloc(sizeof(double) * (size_t)J->np);
  oracle_result res;
  long long N = J->ntraj;
  for (;;) {
    long long i0 = atomic_fetch_add(&J->next, ENS_CHUNK);
    if (i0 >= N) break;
    long long i1 = i0 + ENS_CHUNK < N ? i0 + ENS_CHUNK : N;
    for (long long i = i0; i < i1; i++) {
      int n = J->nvec ? J->nvec[i] : J->n;
      for (int c = 0; c < n; c++) u0[c] = J->u0[(long long)c * N + i];
      mdl_t m;
      if (pl) {
        for (int j = 0; j < J->np; j++) pl[j] = J->p[(long long)j * N + i];
        mdl_init(&m, J->model, n, pl, J->np);
      } else {
        for (int j = 0; j < J->np && j < 16; j++) p[j] = J->p[(long long)j * N + i];
        mdl_init(&m, J->model, n, p, J->np);
      }
      /* value (k, c) of trajectory i: SoA at ((k n + c) N + i), trajectory
       * major at ((i ns + k) n + c), n the layout's rows */
      const int tmaj = J->opt->save_layout == 1;
      double *sv = !J->out_save ? NULL
                   : tmaj ? J->out_save + (long long)i * J->opt->ns * J->n
                          : J->out_save + i;
      solve_one(&m, u0, J->t0[i], J->tf[i], J->opt, (int64_t)i, sv, tmaj ? 1 : N, J->n, &res,
                ws);
      if (J->u_final)
        for (int c = 0; c < n; c++) J->u_final[(long long)c * N + i] = res.u[c];
      if (J->t_final) J->t_final[i] = res.t;
      if (J->retcode) J->retcode[i] = (int8_t)res.retcode;
      if (J->naccept) J->naccept[i] = res.naccept;
      if (J->nreject) J->nreject[i] = res.nreject;
      if (J->nf) J->nf[i] = res.nf;
      if (J->dt0) J->dt0[i] = res.dt0;
      if (J->nstiff) J->nstiff[i] = res.nstiff_steps;
      if (J->dt_first) J->dt_first[i] = res.dt_first_accept;
      if (J->nevent) J->nevent[i] = res.nevent;
      if (J->t_event) J->t_event[i] = res.t_event;
      if (J->opt->ev_n) J->opt->ev_n[i] = res.nevsave;
    }
  }
  free(pl);
  free(ws);
  return NULL;
}

int oracle_ensemble_solve(int32_t model, int32_t n, int32_t np, int64_t ntraj,
                          const double *u0, const double *p, const double *t0,
                          const double *tf, const int32_t *nvec,
                          const oracle_options *opt, int32_t nthreads,
                          double *out_save, double *u_final, double *t_final,
                          int8_t *retcode, int32_t *naccept, int32_t *nreject,
                          int32_t *nf, double *dt0, int32_t *nstiff, double *dt_first,
                          int32_t *nevent, double *t_event) {
  if (n < 1 || n > NMAX || ntraj < 0) return -1;
  ens_job J;
  memset(&J, 0, sizeof(J));
  J.model = model;
  J.n = n;
  J.np = np;
  J.ntraj = ntraj;
  J.u0 = u0;
  J.p = p;
  J.t0 = t0;
  J.tf = tf;
  J.nvec = nvec;
  J.opt = opt;
  J.out_save = out_save;
  J.u_final = u_final;
  J.t_final = t_final;
  J.retcode = retcode;
  J.naccept = naccept;
  J.nreject = nreject;
  J.nf = nf;
  J.dt0 = dt0;
  J.nstiff = nstiff;
  J.dt_first = dt_first;
  J.nevent = nevent;
  J.t_event = t_event;
  atomic_init(&J.next, 0);
  if (nthreads < 1) nthreads = 1;
  if (nthreads == 1 || ntraj <= 1) {
    ens_worker(&J);
    return 0;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, 8u << 20);
  int started = 0;
  for (int i = 0; i < nthreads; i++) {
    if (pthread_create(&th[i], &attr, ens_worker, &J) == 0) started++;
    else break;
  }
  if (started == 0) ens_worker(&J);
  for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
  pthread_attr_destroy(&attr);
  free(th);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Test hooks                                                                */
/* ------------------------------------------------------------------------ */
int oracle_tsit5_tableau(double *c7, double *a49, double *btilde7, double *r28) {
  for (int i = 0; i < 7; i++) c7[i] = TS_C[i];
  for (int i = 0; i < 7; i++)
    for (int j = 0; j < 7; j++) a49[i * 7 + j] = TS_A[i][j];
  for (int i = 0; i < 7; i++) btilde7[i] = TS_BT[i];
  for (int i = 0; i < 7; i++)
    for (int j = 0; j < 4; j++) r28[i * 4 + j] = TS_R[i][j];
  return 0;
}

int oracle_rosenbrock_tableau(int32_t which, double *gamma, double *a64, double *C64,
                              double *c8, double *d8) {
  /* full 8x8 a (stages 7, 8 expanded: a7j = a6j, a76 = 1; a8j = a7j, a87 = 1) */
  const rdtab_t *T = which == 1 ? &RD5P : &RD5;
  *gamma = T->gamma;
  memset(a64, 0, sizeof(double) * 64);
  memset(C64, 0, sizeof(double) * 64);
  for (int i = 1; i < 6; i++)
    for (int j = 0; j < i; j++) a64[i * 8 + j] = T->A[i][j];
  for (int j = 0; j < 5; j++) {
    a64[6 * 8 + j] = T->A[5][j];
    a64[7 * 8 + j] = T->A[5][j];
  }
  a64[6 * 8 + 5] = 1.0;
  a64[7 * 8 + 5] = 1.0;
  a64[7 * 8 + 6] = 1.0;
  for (int i = 1; i < 8; i++)
    for (int j = 0; j < i; j++) C64[i * 8 + j] = T->C[i][j];
  for (int i = 0; i < 8; i++) {
    c8[i] = T->c[i];
    d8[i] = T->d[i];
  }
  return 0;
}

int oracle_rodas5_tableau(double *gamma, double *a64, double *C64, double *c8, double *d8) {
  return oracle_rosenbrock_tableau(0, gamma, a64, C64, c8, d8);
}

void oracle_rhs(int32_t model, int32_t n, const double *u, const double *p,
                double t, double *du) {
  mdl_t m;
  mdl_init(&m, model, n, p, model_np(model, n));
  rhs64(&m, u, t, du);
}

void oracle_jac(int32_t model, int32_t n, const double *u, const double *p,
                double t, double *J) {
  mdl_t m;
  mdl_init(&m, model, n, p, model_np(model, n));
  jac64(&m, u, t, J);
}
