// lane_solver.hpp -- one trajectory per lane: the integrator loop of
// OrdinaryDiffEq (loopheader!/perform_step!/loopfooter!) for Tsit5, Rodas5
// and AutoTsit5(Rodas5()), held entirely in VGPRs.
//
// Reference behaviour restated here: Tsit5 / Rodas5 perform_step, the PI
// controller, ode_determine_initdt, tstop clamping at tf and savevalues!
// for a vector saveat (OrdinaryDiffEq 7, re-exported by
// src/DifferentialEquations.jl:8-9; surface names test/qa/qa.jl:15,97,119,
// 135,185; saveat semantics test/core.jl:90-96, u0 copy core.jl:34).
#pragma once

#include "device_common.hpp"

namespace mideq {

// AUTO: AutoTsit5(Rodas5()); AUTO5P: AutoTsit5(Rodas5P()).  AUTO_T (internal):
// first pass of a two-pass AutoSwitch solve -- Tsit5 with the stiffness
// test, handing every trajectory that would switch (or starts stiff) to a
// second, full AutoSwitch pass; a trajectory that never switches takes the
// same steps either way, so results equal the single-pass solve.
// TRUST (internal): first pass of the predictor-trusting AutoSwitch solve --
// plain Tsit5 (no stiffness test) for the trajectories hinted non-stiff,
// handing to the full AutoSwitch pass every trajectory hinted stiff and every
// one that exceeds its iteration cap (the predictor was wrong about it).
enum { ALG_TSIT5 = 0, ALG_RODAS5 = 1, ALG_AUTO = 2, ALG_RODAS5P = 3, ALG_AUTO5P = 4,
       ALG_AUTO_T = 5, ALG_TRUST = 6 };
// Template flag on a Rosenbrock-capable ALG: build the Rosenbrock dense
// output for saveat.  It keeps the 8 stage increments live past the step,
// so kernels without saveat are built without it.
enum { ALG_DENSE = 16 };

struct LaneArgs {
  int64_t ntraj;
  const double* u0;
  const double* p;
  const double* t0;
  const double* tf;
  double tspan0, tspan1;
  double abstol, reltol;
  double dt_fixed, dtmin, dtmax;
  int64_t maxiters;
  const double* saveat;
  int32_t ns;
  int32_t save_f32;
  const double* dt0;
  const int8_t* alg_hint;
  int32_t qsteady_default;
  int32_t save_layout;  // saved values [ns][N][ntraj] (0) or [ntraj][ns][N] (1)
  void* us;
  double* u_final;
  double* t_final;
  int8_t* retcode;
  int32_t* naccept;
  int32_t* nreject;
  int32_t* nf;
  int32_t* nstiff;
  double* dt_first;
  int32_t* nevent;            // callbacks: affects applied per trajectory
  double* t_event;            // callbacks: time of the last affect
  double* out;                // output_func values [nout][ntraj]
  unsigned long long* queue;  // persistent schedule: next position in the order
  const int64_t* order;       // launch order (NULL: natural order)
  const unsigned long long* order_len;  // device length of `order` (NULL: ntraj)
  int64_t* handoff;           // two-phase AutoSwitch: trajectories for the second pass
  unsigned long long* handoff_count;
  const int32_t* iter_cap;    // TRUST pass: per-trajectory iteration cap (NULL: none)
  double* ev_t;               // callbacks: points saved at events, [ntraj][ev_cap] times,
  double* ev_u;               //   [ntraj][ev_cap][N] states (NULL: none)
  int32_t* ev_n;              //   [ntraj] points produced (may exceed ev_cap)
  int32_t ev_cap;
  int32_t pad1;
  const double* tstops;       // models compiled with tstops: [nts] ascending times every
  int32_t nts;                //   trajectory's steps land on (NULL / 0: tf only)
  int32_t pad2;
};

// Trajectory scheduled at position k of the launch order; -1 for an entry
// outside [0, ntraj) (skipped, so a bad order cannot address out of bounds)
// or past a device-side order length.
__device__ __forceinline__ int64_t lane_traj(const LaneArgs& a, int64_t k) {
  if (a.order_len && (uint64_t)k >= *a.order_len) return -1;
  int64_t i = a.order ? a.order[k] : k;
  return (uint64_t)i < (uint64_t)a.ntraj ? i : -1;
}

// Number of launch-order positions to walk: ntraj, or the device-side order
// length when one is given (the second pass of a two-phase AutoSwitch solve).
__device__ __forceinline__ int64_t lane_queue_len(const LaneArgs& a) {
  if (!a.order_len) return a.ntraj;
  unsigned long long n = *a.order_len;
  return n < (unsigned long long)a.ntraj ? (int64_t)n : a.ntraj;
}

// Append trajectory i to the hand-off list when `mine`: one atomic per
// wavefront (over the lanes active here), lanes placed by ballot prefix.
__device__ __forceinline__ void handoff_append(const LaneArgs& a, int64_t i, bool mine) {
  unsigned long long m = __ballot(mine);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(m);
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(a.handoff_count, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  if (mine) a.handoff[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

// saveat times are staged in LDS by the lane kernels when they fit, so the
// per-step "next save point reached?" test reads a register and a save reads
// LDS: no global load (and no vmcnt wait, which on this ISA also waits for
// the previous step's saveat stores) sits in the step loop.
constexpr int kSaveLds = 2048;
extern __shared__ double lane_saveat_lds[];

// Dynamic LDS bytes a lane kernel launch needs for ns save times.
__host__ __device__ inline size_t lane_lds_bytes(int32_t ns) {
  return (ns > 0 && ns <= kSaveLds) ? (size_t)ns * sizeof(double) : 0;
}

// Whole workgroup: copy the save times into LDS (call before any lane exits).
__device__ __forceinline__ void lane_stage_saveat(const LaneArgs& a) {
  if (a.ns > 0 && a.ns <= kSaveLds) {
    for (int k = threadIdx.x; k < a.ns; k += blockDim.x) lane_saveat_lds[k] = a.saveat[k];
    __syncthreads();
  }
}

template <int N>
__device__ __forceinline__ double rms_sum_small(const double* r) {
  double s = r[0] * r[0];
#pragma unroll
  for (int i = 1; i < N; i++) s = s + r[i] * r[i];
  return s;
}

// calculate_residuals + ODE_DEFAULT_NORM (RMS).
template <int N>
__device__ __forceinline__ double err_norm(const double* ut, const double* u0,
                                           const double* u1, double abstol,
                                           double reltol) {
  double r[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    double sc = fma(fmax(fabs(u0[i]), fabs(u1[i])), reltol, abstol);
    r[i] = ut[i] / sc;
  }
  return sqrt(rms_sum_small<N>(r) / (double)N);
}

// Fast-controller error estimate: EEst^2 from one division over a common
// denominator (n <= 3; an overflowing numerator stays +inf, a rejected
// step), else mean(r^2) without the square root.
template <int N>
__device__ __forceinline__ double err_sq_fast(const double* ut, const double* u0,
                                              const double* u1, double abstol, double reltol) {
  double sc[N];
#pragma unroll
  for (int i = 0; i < N; i++) sc[i] = fma(fmax(fabs(u0[i]), fabs(u1[i])), reltol, abstol);
  if constexpr (N > 3) {
    double r[N];
#pragma unroll
    for (int i = 0; i < N; i++) r[i] = ut[i] / sc[i];
    return rms_sum_small<N>(r) / (double)N;
  }
  double num, den;
  if constexpr (N == 1) {
    num = ut[0] * ut[0];
    den = sc[0] * sc[0];
  } else if constexpr (N == 2) {
    double a0 = ut[0] * sc[1], a1 = ut[1] * sc[0];
    num = a0 * a0;
    num = fma(a1, a1, num);
    double P = sc[0] * sc[1];
    den = (P * P) * 2.0;
  } else {
    static_assert(N == 3 || N > 3, "n >= 1");
    double p01 = sc[0] * sc[1], p02 = sc[0] * sc[2], p12 = sc[1] * sc[2];
    double a0 = ut[0] * p12, a1 = ut[1] * p02, a2 = ut[2] * p01;
    num = a0 * a0;
    num = fma(a1, a1, num);
    num = fma(a2, a2, num);
    double P = p01 * sc[2];
    den = (P * P) * 3.0;
  }
  if (!(num < __builtin_inf())) return num;
  return num / den;
}

template <int N, bool FAST>
__device__ __forceinline__ double err_est(const double* ut, const double* u0, const double* u1,
                                          double abstol, double reltol) {
  if constexpr (FAST)
    return err_sq_fast<N>(ut, u0, u1, abstol, reltol);
  else
    return err_norm<N>(ut, u0, u1, abstol, reltol);
}

template <int N>
__device__ __forceinline__ double rms(const double* x) {
  return sqrt(rms_sum_small<N>(x) / (double)N);
}

// Tsit5 dense-output weights b_i(theta), in the power basis: b_1 = theta +
// r12 theta^2 + r13 theta^3 + r14 theta^4 (r11 = 1), b_i = r_i2 theta^2 +
// r_i3 theta^3 + r_i4 theta^4, lowest power first.  Every operation has at
// most one method constant (an SGPR operand), so a save costs 3 + 21 FP64
// instructions for the weights; Horner's innermost fma(theta, r_i4, r_i3)
// has two and costs two extra VGPR moves per constant.
__device__ __forceinline__ void tsit5_bthetas(double th, double* b) {
  using namespace ts;
  static_assert(r11 == 1.0, "b_1 starts at theta");
  double th2 = th * th;
  double th3 = th2 * th;
  double th4 = th2 * th2;
  b[0] = fma(r14, th4, fma(r13, th3, fma(r12, th2, th)));
  b[1] = fma(r24, th4, fma(r23, th3, r22 * th2));
  b[2] = fma(r34, th4, fma(r33, th3, r32 * th2));
  b[3] = fma(r44, th4, fma(r43, th3, r42 * th2));
  b[4] = fma(r54, th4, fma(r53, th3, r52 * th2));
  b[5] = fma(r64, th4, fma(r63, th3, r62 * th2));
  b[6] = fma(r74, th4, fma(r73, th3, r72 * th2));
}

// AutoSwitch stiffness estimate max_c |dk_c / dg_c| over the components.
// Reference form: one division per component.  Fast form: the running
// maximum kept as a fraction (num, den) and compared by cross-multiplying,
// so no division; a component whose quotient would be NaN (0/0, inf/inf,
// NaN) makes the estimate NaN in both forms.
template <int N, bool FAST>
__device__ __forceinline__ void stiff_estimate(const double* dk, const double* dg, double& num,
                                               double& den) {
  if constexpr (!FAST) {
    double e = 0.0;
    bool nan_seen = false;
#pragma unroll
    for (int c = 0; c < N; c++) {
      double r = fabs(dk[c] / dg[c]);
      nan_seen = nan_seen || (r != r);
      e = (r > e) ? r : e;
    }
    num = nan_seen ? __builtin_nan("") : e;
    den = 1.0;
  } else {
    double en = 0.0, ed = 1.0;
    bool nan_seen = false;
#pragma unroll
    for (int c = 0; c < N; c++) {
      double nn = fabs(dk[c]), dd = fabs(dg[c]);
      nan_seen = nan_seen || (nn != nn) || (dd != dd) || (nn == 0.0 && dd == 0.0) ||
                 (nn == __builtin_inf() && dd == __builtin_inf());
      bool take = nn * ed > en * dd;
      en = take ? nn : en;
      ed = take ? dd : ed;
    }
    num = nan_seen ? __builtin_nan("") : en;
    den = ed;
  }
}

// Callbacks of a model (run-time compiled user models may carry them):
// M::NCB, and per callback k the constexpr kind (0 continuous, 1 discrete),
// direction (continuous: 0 both, 1 up-crossings, -1 down-crossings),
// terminate flag and interpolation points, the condition M::cb_g(k, u, p, t)
// (continuous: a function whose sign changes; discrete: fires when nonzero)
// and the affect M::cb_affect(k, u, p, t) on the state.
template <class M, class = void>
struct cb_count {
  static constexpr int value = 0;
};
template <class M>
struct cb_count<M, decltype((void)M::NCB, void())> {
  static constexpr int value = M::NCB;
};

// On-device output_func of a model: M::NOUT values out[j] of the final
// state, M::output(out, u, p, t, rc).
template <class M, class = void>
struct out_count {
  static constexpr int value = 0;
};
template <class M>
struct out_count<M, decltype((void)M::NOUT, void())> {
  static constexpr int value = M::NOUT;
};

// Non-autonomous right-hand side u' = f(u, p, t): a model with TDEP = true
// takes the time as a last argument of f and jac and gives tgrad(dT, u, p, t)
// = df/dt; the built-in models are autonomous.
template <class M, class = void>
struct time_dep {
  static constexpr bool value = false;
};
template <class M>
struct time_dep<M, decltype((void)M::TDEP, void())> {
  static constexpr bool value = M::TDEP;
};

// save_everystep of a model (M::EVERYSTEP = true): every accepted step's end
// point (t, u) is saved, into the same per-trajectory arrays as the points
// saved at callback events.
template <class M, class = void>
struct every_step {
  static constexpr bool value = false;
};
template <class M>
struct every_step<M, decltype((void)M::EVERYSTEP, void())> {
  static constexpr bool value = M::EVERYSTEP;
};

// tstops of a model (M::TSTOPS = true): the steps land on the times of
// LaneArgs.tstops inside the span, as on tf.
template <class M, class = void>
struct has_tstops {
  static constexpr bool value = false;
};
template <class M>
struct has_tstops<M, decltype((void)M::TSTOPS, void())> {
  static constexpr bool value = M::TSTOPS;
};

template <class M, bool MIX, int ALG, bool FAST>
struct Lane {
  static constexpr int N = M::N;
  static constexpr bool kTdep = time_dep<M>::value;
  static constexpr int kNcb = cb_count<M>::value;
  static constexpr bool kEveryStep = every_step<M>::value;
  static constexpr bool kTstops = has_tstops<M>::value;
  // points saved beside saveat (at events, at every step): kept only by
  // the kernels of models that make them
  static constexpr bool kExtra = kNcb > 0 || kEveryStep;
  static constexpr int kNout = out_count<M>::value;
  static constexpr int NP = M::NP;
  static constexpr int kAlg = ALG & (ALG_DENSE - 1);
  static constexpr bool kDense = (ALG & ALG_DENSE) != 0;
  static constexpr bool kAuto = (kAlg == ALG_AUTO || kAlg == ALG_AUTO5P || kAlg == ALG_AUTO_T);
  static constexpr bool kHasTsit5 = (kAlg != ALG_RODAS5 && kAlg != ALG_RODAS5P);
  static constexpr bool kHasRodas = (kAlg != ALG_TSIT5 && kAlg != ALG_AUTO_T && kAlg != ALG_TRUST);
  static constexpr bool kTrust = (kAlg == ALG_TRUST);
  static constexpr bool kHandoff = (kAlg == ALG_AUTO_T || kTrust);
  static constexpr bool kP = (kAlg == ALG_RODAS5P || kAlg == ALG_AUTO5P);  // Rodas5P tableau
  // the division-free stiffness test of the fast controller (n <= 3)
  static constexpr bool kFastStiff = FAST && N <= 3;

  int64_t i;
  double p[NP];
  float pf[NP];
  double u[N];
  double k1[N];   // FSAL stage (FP64 form)
  float k1f[N];   // FSAL stage (FP32 form)
  double t, tcomp, tf, dt, dtmax;
  bool back;    // tf < t0: integrated backward with steps of -dt (dt > 0 is the magnitude)
  double tdir;  // +1 or -1 (signs by multiplication: one instruction, no selects)
  double tsnap;  // 100 eps(tf): a forward step ending this close to tf ends at tf
  double qold, qoldpow, qsteady_max, eig, dtfirst, lqold;
  double eigd;  // fast form: the stiffness estimate is eig / eigd
  int64_t iter;
  int32_t naccept, nreject, nf, nstiff;
  int32_t isave, count, nanrej;
  int8_t rc, mode;
  double tnext;  // saveat[isave]; +inf (-inf backward) once every point is saved
  // callbacks: condition values at the last step end, affects applied, and
  // the time of the last one
  double gprev[kNcb > 0 ? kNcb : 1];
  int32_t nevent;
  double tevent;
  int32_t nevsave;  // points saved at events (and every step) so far
  double tlast;     // time of the last saved point (saveat or extra), t0 at the start
  int32_t its;      // tstops: index of the next one ahead (-1 / nts: none)

  // The first tstop ahead of t: the next of LaneArgs.tstops strictly inside
  // (t, tf) in the direction of integration, else tf (upstream keeps tf in
  // the same queue).
  __device__ __forceinline__ double next_tstop(const LaneArgs& a) const {
    if constexpr (kTstops) {
      if (its >= 0 && its < a.nts) {
        const double s = a.tstops[its];
        if (tdir * (tf - s) > 0.0) return s;
      }
    }
    return tf;
  }
  // handle_tstop!: drop the tstops at or behind t
  __device__ __forceinline__ void pop_tstops(const LaneArgs& a) {
    if constexpr (kTstops) {
      while (its >= 0 && its < a.nts && tdir * (t - a.tstops[its]) >= 0.0) its += tdir > 0.0 ? 1 : -1;
    }
  }
  int32_t cap;  // TRUST pass: iteration cap of this trajectory

  // point k of this trajectory's event points: (t, v), stored while k < ev_cap
  __device__ __forceinline__ void save_event(const LaneArgs& a, const double* v) {
    if (!a.ev_t) return;
    if (nevsave < a.ev_cap) {
      const int64_t at = (int64_t)i * a.ev_cap + nevsave;
      a.ev_t[at] = t;
#pragma unroll
      for (int c = 0; c < N; c++) a.ev_u[at * N + c] = v[c];
    }
    nevsave++;
  }

  // ------------------------------------------------------------ save times
  __device__ __forceinline__ double save_time(const LaneArgs& a, int k) const {
    if (a.ns <= kSaveLds) return lane_saveat_lds[k];
    return ((const __attribute__((address_space(1))) double*)a.saveat)[k];
  }
  __device__ __forceinline__ void set_tnext(const LaneArgs& a) {
    // save points in ascending order, visited from the last one down when
    // integrating backward; +-inf once every point is saved
    if (back)
      tnext = isave >= 0 ? save_time(a, isave) : -__builtin_inf();
    else
      tnext = isave < a.ns ? save_time(a, isave) : __builtin_inf();
  }

  // ------------------------------------------------------------ helpers
  __device__ __forceinline__ void store_save(const LaneArgs& a, int k, const double* v) {
    if (a.save_layout) {
      // element (i, k, 0) of the [ntraj][ns][N] output: the point's N values
      // are contiguous, one partial line per lane and point; ns * N < 2^32
      // and ntraj < 2^32 (checked by the host)
      const int64_t o =
          (int64_t)((uint64_t)(uint32_t)i * (uint32_t)(a.ns * N) + (uint64_t)(uint32_t)(k * N));
      if (a.save_f32) {
        float* out = (float*)a.us + o;
#pragma unroll
        for (int c = 0; c < N; c++) out[c] = (float)v[c];
      } else {
        double* out = (double*)a.us + o;
#pragma unroll
        for (int c = 0; c < N; c++) out[c] = v[c];
      }
      return;
    }
    // element (k, 0, i) of the [ns][N][ntraj] output; N * ntraj < 2^32 (checked
    // by the host), so the offset is one 32 x 32 -> 64-bit multiply-add
    const int64_t NT = a.ntraj;
    const int64_t o = (int64_t)((uint64_t)(uint32_t)k * (uint32_t)(N * NT) + (uint64_t)i);
    if (a.save_f32) {
      float* out = (float*)a.us + o;
#pragma unroll
      for (int c = 0; c < N; c++) out[c * NT] = (float)v[c];
    } else {
      double* out = (double*)a.us + o;
#pragma unroll
      for (int c = 0; c < N; c++) out[c * NT] = v[c];
    }
  }

  // f at time tt (ignored by autonomous models); T is double, float or a
  // dual number, the time is passed in that type
  template <class T>
  __device__ __forceinline__ void rhs(T* du, const T* uu, const T* pp, double tt) const {
    if constexpr (kTdep)
      M::f(du, uu, pp, (T)tt);
    else
      M::f(du, uu, pp);
  }

  __device__ __forceinline__ void fsal_init() {
    rhs(k1, u, p, t);
    if (MIX) {
      float uf[N];
#pragma unroll
      for (int c = 0; c < N; c++) uf[c] = (float)u[c];
      rhs(k1f, uf, pf, t);
    }
  }

  // ode_determine_initdt (Hairer-Norsett-Wanner II.4 with upstream edits).
  __device__ double initdt(const LaneArgs& a, double t0) {
    double dtmin_e = nextafter(fmax(a.dtmin, jl_eps(t0)), __builtin_inf());
    double smalldt = fmax(dtmin_e, 1e-6);
    double sk[N], tmp[N];
#pragma unroll
    for (int c = 0; c < N; c++) sk[c] = fma(fabs(u[c]), a.reltol, a.abstol);
#pragma unroll
    for (int c = 0; c < N; c++) tmp[c] = u[c] / sk[c];
    double d0 = rms<N>(tmp);
#pragma unroll
    for (int c = 0; c < N; c++) tmp[c] = k1[c] / sk[c];
    double d1 = rms<N>(tmp);
    if (d1 != d1) return dtmin_e;
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : (d0 / d1) / 100.0;
    dt0 = fmin(dt0, dtmax);
    if (dt0 < 10.0 * 2.220446049250313e-16) return smalldt;
    double u1[N], f1[N];
    double sdt0 = tdir * dt0;  // the trial step in the direction of integration
#pragma unroll
    for (int c = 0; c < N; c++) u1[c] = fma(sdt0, k1[c], u[c]);
    rhs(f1, u1, p, t0 + sdt0);
    nf += 1;
    bool same = true;
#pragma unroll
    for (int c = 0; c < N; c++) same = same && (k1[c] == f1[c]);
    if (same) return fmax(dtmin_e, 100.0 * dt0);
#pragma unroll
    for (int c = 0; c < N; c++) tmp[c] = (f1[c] - k1[c]) / sk[c];
    double d2 = rms<N>(tmp) / dt0;
    double mx = fmax(d1, d2);
    double dt1;
    if (mx <= 1e-15)
      dt1 = fmax(1e-6, dt0 * 1e-3);
    else
      dt1 = ctl_exp2(-(2.0 * 3.32192809488736234787 + ctl_log2(mx)) / 5.0);
    return fmax(dtmin_e, fmin(fmin(100.0 * dt0, dt1), dtmax));
  }

  // ------------------------------------------------------------ init
  __device__ void init(const LaneArgs& a, int64_t idx) {
    i = idx;
    const int64_t NT = a.ntraj;
#pragma unroll
    for (int j = 0; j < NP; j++) {
      p[j] = a.p[(int64_t)j * NT + i];
      pf[j] = (float)p[j];
    }
#pragma unroll
    for (int c = 0; c < N; c++) u[c] = a.u0[(int64_t)c * NT + i];
    double t0 = a.t0 ? a.t0[i] : a.tspan0;
    tf = a.tf ? a.tf[i] : a.tspan1;
    back = tf < t0;
    tdir = back ? -1.0 : 1.0;
    t = t0;
    tcomp = 0.0;
    dtmax = a.dtmax > 0 ? a.dtmax : fabs(tf - t0);
    // a step ending within 100 eps(max(t, tf)) of tf ends at tf, t the
    // step's start: forward t < tf in the loop, so the window is fixed;
    // backward max(t, tf) = t and the window is taken per step
    tsnap = 100.0 * jl_eps(tf);
    naccept = nreject = nstiff = 0;
    nf = 1;
    iter = 0;
    count = 0;
    nanrej = 0;
    eig = 0.0;
    eigd = 1.0;
    dtfirst = 0.0;
    rc = RC_DEFAULT;
    mode = (kAlg == ALG_RODAS5 || kAlg == ALG_RODAS5P) ? 1 : 0;
    if (kAuto && a.alg_hint) mode = a.alg_hint[i] ? 1 : 0;
    if (kHandoff && mode == 1) {
      rc = RC_HANDOFF;  // starts stiff: the full AutoSwitch pass solves it
      return;
    }
    if constexpr (kTrust) {
      if (a.alg_hint && a.alg_hint[i]) {
        rc = RC_HANDOFF;  // predicted stiff: the AutoSwitch pass solves it
        return;
      }
      cap = a.iter_cap ? a.iter_cap[i] : 0x7fffffff;
    }
    qsteady_max = (a.qsteady_default && mode == 1) ? 1.2 : 1.0;
    // saveat points before t0 are unreachable; a point at t0 is u0 itself
    isave = 0;
    if (a.us) {
      double nanv[N];
#pragma unroll
      for (int c = 0; c < N; c++) nanv[c] = __builtin_nan("");
      if (!back) {
        while (isave < a.ns && save_time(a, isave) < t0) store_save(a, isave++, nanv);
        while (isave < a.ns && save_time(a, isave) == t0) store_save(a, isave++, u);
      } else {
        isave = a.ns - 1;
        while (isave >= 0 && save_time(a, isave) > t0) store_save(a, isave--, nanv);
        while (isave >= 0 && save_time(a, isave) == t0) store_save(a, isave--, u);
      }
    } else {
      if (!back) {
        while (isave < a.ns && save_time(a, isave) <= t0) isave++;
      } else {
        isave = a.ns - 1;
        while (isave >= 0 && save_time(a, isave) >= t0) isave--;
      }
    }
    set_tnext(a);
    fsal_init();
    nevent = 0;
    tevent = __builtin_nan("");
    nevsave = 0;
    tlast = t0;
    if constexpr (kTstops) {
      // the first tstop strictly after t0 (forward) or before it (backward)
      its = tdir > 0.0 ? 0 : a.nts - 1;
      pop_tstops(a);
    }
    if constexpr (kNcb > 0) {
#pragma unroll
      for (int k = 0; k < kNcb; k++) gprev[k] = M::cb_g(k, u, p, t0);
    }
    if (a.dt_fixed > 0)
      dt = a.dt_fixed;
    else if (a.dt0)
      dt = a.dt0[i];
    else
      dt = initdt(a, t0);
    qold = kQoldInit;
    qoldpow = ctl_pow(kQoldInit, kBeta2);
    lqold = kL2QoldInit;
    // a zero-length span is done; a NaN end is not integrated
    if (tf == t0)
      rc = RC_SUCCESS;
    else if (!(tf > t0) && !(tf < t0))
      rc = RC_INITIALFAILURE;
  }

  // ------------------------------------------------------------ Tsit5
  __device__ double tsit5_step64(double dts, double k[7][N], double* u1, double abstol,
                                 double reltol) {
    using namespace ts;
    double g[N], g6[N], ut[N];
    double aa = dts * a21;
#pragma unroll
    for (int c = 0; c < N; c++) {
      k[0][c] = k1[c];
      g[c] = fma(aa, k[0][c], u[c]);
    }
    rhs(k[1], g, p, t + ts::c2 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = a32 * k[1][c];
      acc = fma(a31, k[0][c], acc);
      g[c] = fma(dts, acc, u[c]);
    }
    rhs(k[2], g, p, t + ts::c3 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = a43 * k[2][c];
      acc = fma(a42, k[1][c], acc);
      acc = fma(a41, k[0][c], acc);
      g[c] = fma(dts, acc, u[c]);
    }
    rhs(k[3], g, p, t + ts::c4 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = a54 * k[3][c];
      acc = fma(a53, k[2][c], acc);
      acc = fma(a52, k[1][c], acc);
      acc = fma(a51, k[0][c], acc);
      g[c] = fma(dts, acc, u[c]);
    }
    rhs(k[4], g, p, t + ts::c5 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = a65 * k[4][c];
      acc = fma(a64, k[3][c], acc);
      acc = fma(a63, k[2][c], acc);
      acc = fma(a62, k[1][c], acc);
      acc = fma(a61, k[0][c], acc);
      g6[c] = fma(dts, acc, u[c]);
    }
    rhs(k[5], g6, p, t + dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = a76 * k[5][c];
      acc = fma(a75, k[4][c], acc);
      acc = fma(a74, k[3][c], acc);
      acc = fma(a73, k[2][c], acc);
      acc = fma(a72, k[1][c], acc);
      acc = fma(a71, k[0][c], acc);
      u1[c] = fma(dts, acc, u[c]);
    }
    rhs(k[6], u1, p, t + dts);
    if (kAuto) {
      double dk[N], dg[N];
#pragma unroll
      for (int c = 0; c < N; c++) {
        dk[c] = k[6][c] - k[5][c];
        dg[c] = u1[c] - g6[c];
      }
      stiff_estimate<N, kFastStiff>(dk, dg, eig, eigd);
    }
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = bt7 * k[6][c];
      acc = fma(bt6, k[5][c], acc);
      acc = fma(bt5, k[4][c], acc);
      acc = fma(bt4, k[3][c], acc);
      acc = fma(bt3, k[2][c], acc);
      acc = fma(bt2, k[1][c], acc);
      acc = fma(bt1, k[0][c], acc);
      ut[c] = dts * acc;
    }
    return err_est<N, FAST>(ut, u, u1, abstol, reltol);
  }

  __device__ double tsit5_step_mix(double dts, float k[7][N], double* u1, double abstol,
                                   double reltol) {
    using namespace ts;
    float u32[N], g[N], g6[N], u1f[N];
    double ut[N];
    float d32 = (float)dts;
#pragma unroll
    for (int c = 0; c < N; c++) {
      u32[c] = (float)u[c];
      k[0][c] = k1f[c];
    }
    float aa = d32 * (float)a21;
#pragma unroll
    for (int c = 0; c < N; c++) g[c] = fmaf(aa, k[0][c], u32[c]);
    rhs(k[1], g, pf, t + ts::c2 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      float acc = (float)a32 * k[1][c];
      acc = fmaf((float)a31, k[0][c], acc);
      g[c] = fmaf(d32, acc, u32[c]);
    }
    rhs(k[2], g, pf, t + ts::c3 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      float acc = (float)a43 * k[2][c];
      acc = fmaf((float)a42, k[1][c], acc);
      acc = fmaf((float)a41, k[0][c], acc);
      g[c] = fmaf(d32, acc, u32[c]);
    }
    rhs(k[3], g, pf, t + ts::c4 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      float acc = (float)a54 * k[3][c];
      acc = fmaf((float)a53, k[2][c], acc);
      acc = fmaf((float)a52, k[1][c], acc);
      acc = fmaf((float)a51, k[0][c], acc);
      g[c] = fmaf(d32, acc, u32[c]);
    }
    rhs(k[4], g, pf, t + ts::c5 * dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      float acc = (float)a65 * k[4][c];
      acc = fmaf((float)a64, k[3][c], acc);
      acc = fmaf((float)a63, k[2][c], acc);
      acc = fmaf((float)a62, k[1][c], acc);
      acc = fmaf((float)a61, k[0][c], acc);
      g6[c] = fmaf(d32, acc, u32[c]);
    }
    rhs(k[5], g6, pf, t + dts);
#pragma unroll
    for (int c = 0; c < N; c++) {
      float acc = (float)a76 * k[5][c];
      acc = fmaf((float)a75, k[4][c], acc);
      acc = fmaf((float)a74, k[3][c], acc);
      acc = fmaf((float)a73, k[2][c], acc);
      acc = fmaf((float)a72, k[1][c], acc);
      acc = fmaf((float)a71, k[0][c], acc);
      u1[c] = fma(dts, (double)acc, u[c]);
      u1f[c] = (float)u1[c];
    }
    rhs(k[6], u1f, pf, t + dts);
    if (kAuto) {
      double dk[N], dg[N];
#pragma unroll
      for (int c = 0; c < N; c++) {
        dk[c] = (double)k[6][c] - (double)k[5][c];
        dg[c] = u1[c] - (double)g6[c];
      }
      stiff_estimate<N, kFastStiff>(dk, dg, eig, eigd);
    }
#pragma unroll
    for (int c = 0; c < N; c++) {
      float acc = (float)bt7 * k[6][c];
      acc = fmaf((float)bt6, k[5][c], acc);
      acc = fmaf((float)bt5, k[4][c], acc);
      acc = fmaf((float)bt4, k[3][c], acc);
      acc = fmaf((float)bt3, k[2][c], acc);
      acc = fmaf((float)bt2, k[1][c], acc);
      acc = fmaf((float)bt1, k[0][c], acc);
      ut[c] = dts * (double)acc;
    }
    return err_est<N, FAST>(ut, u, u1, abstol, reltol);
  }

  // ------------------------------------------------------------ Rodas5
  // W = J - I/(gamma dt), LU with partial pivoting in registers (row swaps
  // as selects so that every index stays static), 8 stage solves.
  __device__ bool rodas5_step(double dts, double* u1, double& EEst, double abstol,
                              double reltol, double (&kd)[8][N]) {
    double W[N][N];
    if constexpr (kTdep)
      M::jac(W, u, p, t);
    else
      M::jac(W, u, p);
    if (kAuto) {
      double best = 0.0;
#pragma unroll
      for (int r = 0; r < N; r++) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < N; c++) s += fabs(W[r][c]);
        best = (s > best) ? s : best;
      }
      eig = best;
      eigd = 1.0;
    }
    double dtgamma = dts * rd_gamma<kP>();
    double w = 1.0 / dtgamma;
    double inv_dt = 1.0 / dts;
#pragma unroll
    for (int r = 0; r < N; r++) W[r][r] = W[r][r] - w;
    int piv[N];
    double rpiv[N];
    bool singular = false;
#pragma unroll
    for (int kk = 0; kk < N; kk++) {
      int pk = kk;
      double best = fabs(W[kk][kk]);
#pragma unroll
      for (int r = kk + 1; r < N; r++) {
        double v = fabs(W[r][kk]);
        if (v > best) {
          best = v;
          pk = r;
        }
      }
      piv[kk] = pk;
#pragma unroll
      for (int r = kk + 1; r < N; r++) {
        bool sw = (pk == r);
#pragma unroll
        for (int c = 0; c < N; c++) {
          double x = W[kk][c], y = W[r][c];
          W[kk][c] = sw ? y : x;
          W[r][c] = sw ? x : y;
        }
      }
      double d = W[kk][kk];
      singular = singular || (d == 0.0);
      double rr = 1.0 / d;
      rpiv[kk] = rr;
#pragma unroll
      for (int r = kk + 1; r < N; r++) {
        double l = W[r][kk] * rr;
        W[r][kk] = l;
#pragma unroll
        for (int c = kk + 1; c < N; c++) W[r][c] = fma(-l, W[kk][c], W[r][c]);
      }
    }
    if (singular) {
      EEst = 2.0;
      return false;
    }
    double dT[N];  // df/dt at (uprev, t); zero for autonomous models
#pragma unroll
    for (int c = 0; c < N; c++) dT[c] = 0.0;
    if constexpr (kTdep) M::tgrad(dT, u, p, t);
    double U[N], du[N], x[N];
    float Uf[N], duf[N];
#pragma unroll
    for (int s = 0; s < 8; s++) {
      if (!MIX) {
        if (s == 0) {
#pragma unroll
          for (int c = 0; c < N; c++) U[c] = u[c];
        } else if (s <= 5) {
#pragma unroll
          for (int c = 0; c < N; c++) {
            double acc = u[c];
#pragma unroll
            for (int j = 0; j < s; j++) acc = fma(rd_A<kP>(s, j), kd[j][c], acc);
            U[c] = acc;
          }
        } else {
#pragma unroll
          for (int c = 0; c < N; c++) U[c] = U[c] + kd[s - 1][c];
        }
        rhs(du, U, p, t + rd_c<kP>(s) * dts);
      } else {
        if (s == 0) {
#pragma unroll
          for (int c = 0; c < N; c++) Uf[c] = (float)u[c];
        } else if (s <= 5) {
#pragma unroll
          for (int c = 0; c < N; c++) {
            float acc = (float)u[c];
#pragma unroll
            for (int j = 0; j < s; j++) acc = fmaf((float)rd_A<kP>(s, j), (float)kd[j][c], acc);
            Uf[c] = acc;
          }
        } else {
#pragma unroll
          for (int c = 0; c < N; c++) Uf[c] = Uf[c] + (float)kd[s - 1][c];
        }
        rhs(duf, Uf, pf, t + rd_c<kP>(s) * dts);
#pragma unroll
        for (int c = 0; c < N; c++) du[c] = (double)duf[c];
      }
#pragma unroll
      for (int c = 0; c < N; c++) {
        double base = du[c];
        if constexpr (kTdep) base = base + (dts * rd_d<kP>(s)) * dT[c];
        double r;
        if (s == 0) {
          r = base;
        } else if (!MIX) {
          double acc = (rd_C<kP>(s, 0) * inv_dt) * kd[0][c];
#pragma unroll
          for (int j = 1; j < s; j++) acc = fma(rd_C<kP>(s, j) * inv_dt, kd[j][c], acc);
          r = base + acc;
        } else {
          float idt = (float)inv_dt;
          float acc = ((float)rd_C<kP>(s, 0) * idt) * (float)kd[0][c];
#pragma unroll
          for (int j = 1; j < s; j++) acc = fmaf((float)rd_C<kP>(s, j) * idt, (float)kd[j][c], acc);
          r = (double)((float)base + acc);
        }
        x[c] = -r;
      }
      // apply the row permutation, then L and U solves (column oriented)
#pragma unroll
      for (int kk = 0; kk < N; kk++) {
#pragma unroll
        for (int r = kk + 1; r < N; r++) {
          bool sw = (piv[kk] == r);
          double a0 = x[kk], b0 = x[r];
          x[kk] = sw ? b0 : a0;
          x[r] = sw ? a0 : b0;
        }
      }
#pragma unroll
      for (int j = 0; j < N; j++)
#pragma unroll
        for (int r = j + 1; r < N; r++) x[r] = fma(-W[r][j], x[j], x[r]);
#pragma unroll
      for (int j = N - 1; j >= 0; j--) {
        x[j] = x[j] * rpiv[j];
#pragma unroll
        for (int r = 0; r < j; r++) x[r] = fma(-W[r][j], x[j], x[r]);
      }
#pragma unroll
      for (int c = 0; c < N; c++) kd[s][c] = MIX ? (double)(float)x[c] : x[c];
    }
#pragma unroll
    for (int c = 0; c < N; c++) {
      double acc = u[c];
#pragma unroll
      for (int j = 0; j < 5; j++) acc = fma(rd_A<kP>(5, j), kd[j][c], acc);
      acc = acc + kd[5][c];
      acc = acc + kd[6][c];
      u1[c] = acc + kd[7][c];
    }
    EEst = err_est<N, FAST>(kd[7], u, u1, abstol, reltol);
    return true;
  }

  // ------------------------------------------------------------ one step
  // One pass of loopheader!/perform_step!/loopfooter!.  Sets rc when the
  // trajectory finishes.
  __device__ void step(const LaneArgs& a) {
    if (kAuto && iter > 0) {
      bool stiff;
      if constexpr (!kFastStiff) {
        double stiffness = fabs(eig * dt / kStabSize);
        stiff = stiffness > kStiffTol;
      } else {
        // |eig dt / S| > tol  <=>  num * dt > (tol S) * den   (num, dt >= 0, den > 0)
        stiff = eig * dt > kStiffTolS * eigd;
      }
      count = stiff ? (count < 0 ? 1 : count + 1) : (count > 0 ? -1 : count - 1);
      if (mode == 0 && count > kMaxStiff) {
        if constexpr (kHandoff) {
          rc = RC_HANDOFF;  // would switch to the Rosenbrock method
          return;
        }
        dt = dt * kDtFac;
        mode = 1;
        qsteady_max = a.qsteady_default ? 1.2 : 1.0;
      } else if (mode == 1 && count < -kMaxNonStiff) {
        dt = dt / kDtFac;
        mode = 0;
        qsteady_max = 1.0;
        fsal_init();
        nf += 1;
      }
    }
    iter++;
    int64_t maxit = a.maxiters > 0 ? a.maxiters : 100000;
    if constexpr (kTrust) {
      if (iter > cap && iter <= maxit) {
        rc = RC_HANDOFF;  // more steps than predicted: hand it to AutoSwitch
        return;
      }
    }
    if (iter > maxit) {
      rc = RC_MAXITERS;
      return;
    }
    dt = fmin(dt, dtmax);
    if (a.dtmin > 0 && dt < a.dtmin) {
      rc = RC_DTLESSTHANMIN;
      return;
    }
    pop_tstops(a);  // handle_tstop!: the ones the last step reached
    [[maybe_unused]] const double tstop = next_tstop(a);
    if constexpr (kTstops)
      dt = fmin(dt, tdir * (tstop - t));  // modify_dt_for_tstops!
    else
      dt = fmin(dt, tdir * (tf - t));
    const double sdt = tdir * dt;  // the signed step

    double u1[N];
    double EEst = 0.0;
    double k64[kHasTsit5 && !MIX ? 7 : 1][N];
    float k32[kHasTsit5 && MIX ? 7 : 1][N];
    double kd[8][N];  // Rosenbrock stage increments (dense output)
    bool tsit_step = kHasTsit5 && (!kHasRodas || mode == 0);
    if (tsit_step) {
      if constexpr (kHasTsit5) {
        if constexpr (!MIX)
          EEst = tsit5_step64(sdt, k64, u1, a.abstol, a.reltol);
        else
          EEst = tsit5_step_mix(sdt, k32, u1, a.abstol, a.reltol);
      }
      nf += 6;
    } else {
      if constexpr (kHasRodas) {
        if (!rodas5_step(sdt, u1, EEst, a.abstol, a.reltol, kd)) EEst = 2.0;
      }
      nf += 8;
    }
    if (EEst != EEst) {
      // FP64: DtNaN as upstream.  Mixed: FP32 stages can overflow on a far
      // too large trial step; treat that as EEst = +inf (reject, dt/5) and
      // give up after 20 consecutive such rejections.
      if (!MIX || ++nanrej > 20) {
        rc = RC_DTNAN;
        return;
      }
      EEst = __builtin_inf();
    } else {
      nanrej = 0;
    }
    // PI controller (the fast form works on EEst^2 in the log2 domain)
    double q = 1.0, q11 = 1.0, lE = 0.0, lq = 0.0, l11 = 0.0, lE2 = 0.0;
    if constexpr (!FAST) {
      if (EEst == 0.0) {
        q = 1.0 / kQmax;
      } else {
        lE = ctl_log2(EEst);
        q11 = ctl_exp2(kBeta1 * lE);
        q = q11 / qoldpow;
        q = fmax(1.0 / kQmax, fmin(1.0 / kQmin, q / kGammaSafety));
      }
    } else {
      if (EEst == 0.0) {
        lq = kL2QmaxInv;
      } else {
        lE2 = fast_log2(EEst);
        l11 = (0.5 * kBeta1) * lE2;
        lq = l11 - kBeta2 * lqold - kL2Gamma;
        lq = fmax(kL2QmaxInv, fmin(kL2QminInv, lq));
      }
    }
    if (EEst <= 1.0) {
      if (naccept == 0) dtfirst = dt;
      naccept++;
      if (!tsit_step) nstiff++;
      double dtnew;
      if constexpr (!FAST) {
        if (1.0 <= q && q <= qsteady_max) q = 1.0;
        qold = fmax(EEst, kQoldInit);
        qoldpow = (qold == EEst) ? ctl_exp2(kBeta2 * lE) : ctl_pow(kQoldInit, kBeta2);
        dtnew = dt / q;
      } else {
        double lqs = qsteady_max > 1.0 ? kL2QsteadyRodas : 0.0;
        if (0.0 <= lq && lq <= lqs) lq = 0.0;
        lqold = (EEst == 0.0) ? kL2QoldInit : fmax(0.5 * lE2, kL2QoldInit);
        dtnew = dt * fast_exp2(-lq);
      }
      double tprev = t;
      double ttmp;
      if (!MIX) {
        ttmp = t + sdt;
      } else {
        double y = sdt - tcomp;
        ttmp = t + y;
        tcomp = (ttmp - t) - y;
      }
      if constexpr (kTstops) {
        // fixed_t_for_floatingpoint_error!: onto the first tstop within
        // 100 eps(max(t, tstop)) (tf included: then the window is the one below)
        if (fabs(ttmp - tstop) < 100.0 * jl_eps(fmax(t, tstop))) ttmp = tstop;
      } else {
        const double snapw = back ? 100.0 * jl_eps(fmax(t, tf)) : tsnap;
        if (fabs(ttmp - tf) < snapw) ttmp = tf;
      }
      t = ttmp;
      // dense output of this step at th (time tprev + th dt): Tsit5's
      // interpolant or the Rosenbrock one (rd_H, FP64 in both precisions)
      auto interp = [&](double th, double* val) {
        if (tsit_step) {
          double bt[7];
          tsit5_bthetas(th, bt);
          if constexpr (kHasTsit5 && !MIX) {
#pragma unroll
            for (int c = 0; c < N; c++) {
              double acc = bt[6] * k64[6][c];
#pragma unroll
              for (int j = 5; j >= 0; j--) acc = fma(bt[j], k64[j][c], acc);
              val[c] = fma(sdt, acc, u[c]);
            }
          } else if constexpr (kHasTsit5 && MIX) {
            float bf[7];
#pragma unroll
            for (int j = 0; j < 7; j++) bf[j] = (float)bt[j];
#pragma unroll
            for (int c = 0; c < N; c++) {
              float acc = bf[6] * k32[6][c];
#pragma unroll
              for (int j = 5; j >= 0; j--) acc = fmaf(bf[j], k32[j][c], acc);
              val[c] = fma(sdt, (double)acc, u[c]);
            }
          }
        } else if constexpr (kHasRodas && !kDense) {
          // not reached: saveat and callbacks run the kDense build
#pragma unroll
          for (int c = 0; c < N; c++) val[c] = __builtin_nan("");
        } else if constexpr (kHasRodas) {
          double th1 = 1.0 - th;
#pragma unroll
          for (int c = 0; c < N; c++) {
            double K[3];
#pragma unroll
            for (int m = 0; m < 3; m++) {
              double acc = rd_H<kP>(m, 0) * kd[0][c];
#pragma unroll
              for (int j = 1; j < 7; j++) acc = fma(rd_H<kP>(m, j), kd[j][c], acc);
              K[m] = acc;
            }
            double inner = fma(th, K[2], K[1]);
            inner = fma(th, inner, K[0]);
            double tt = fma(th1, inner, u1[c]);
            val[c] = fma(th, tt, th1 * u[c]);
          }
        }
      };
      // the state the step ends in: u1, or the dense output at an event
      double uend[N];
#pragma unroll
      for (int c = 0; c < N; c++) uend[c] = u1[c];
      int ev = -1;
      if constexpr (kNcb > 0) {
        static_assert(!kHasRodas || kDense, "callbacks need the dense output");
        // continuous callbacks: the earliest sign change of any condition
        // over interp_points sub-intervals of the step, located by the
        // Illinois (modified regula falsi) bracketing iteration on th; the
        // event is the bracket end past the crossing
        auto G = [&](int k, double th) {
          double tmp[N];
          interp(th, tmp);
          return M::cb_g(k, tmp, p, tprev + th * sdt);
        };
        double th_ev = 2.0;
        for (int k = 0; k < kNcb; k++) {
          if (M::cb_kind(k) != 0) continue;
          const int ip = M::cb_interp(k);
          const int dir = M::cb_dir(k);
          double ga = gprev[k], tha = 0.0;
          for (int q = 1; q <= ip; q++) {
            double thb, gb;
            if (q == ip) {
              thb = 1.0;
              gb = M::cb_g(k, u1, p, t);
            } else {
              thb = (double)q / (double)ip;
              gb = G(k, thb);
            }
            bool up = ga < 0.0 && gb >= 0.0;
            bool down = ga > 0.0 && gb <= 0.0;
            if ((up && dir >= 0) || (down && dir <= 0)) {
              double th = thb;
              if (gb != 0.0) {
                double lo = tha, hi = thb, flo = ga, fhi = gb;
                int side = 0;
                for (int it = 0; it < 100; it++) {
                  if (hi - lo <= 4.0 * 2.220446049250313e-16) break;
                  double c = (lo * fhi - hi * flo) / (fhi - flo);
                  if (!(c > lo && c < hi)) c = 0.5 * (lo + hi);
                  double fc = G(k, c);
                  if (fc == 0.0) {
                    hi = c;
                    break;
                  }
                  if ((fc < 0.0) == (fhi < 0.0)) {
                    hi = c;
                    fhi = fc;
                    if (side == -1) flo = 0.5 * flo;
                    side = -1;
                  } else {
                    lo = c;
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
          interp(th_ev, uend);
        }
      }
      // savevalues!: the saveat points up to t (exact: one of them is t)
      [[maybe_unused]] bool exact = false;
      {
        while (tdir * (t - tnext) >= 0.0) {  // tnext at or behind t (+-inf: never)
          if (a.us) {
            double val[N];
            double sa = tnext;
            if (sa == t) {
#pragma unroll
              for (int c = 0; c < N; c++) val[c] = uend[c];
            } else {
              interp((sa - tprev) / sdt, val);
            }
            store_save(a, isave, val);
          }
          if constexpr (kExtra) {
            exact = exact || tnext == t;
            tlast = tnext;
          }
          isave += back ? -1 : 1;
          set_tnext(a);
        }
      }
#pragma unroll
      for (int c = 0; c < N; c++) u[c] = uend[c];
      if (tsit_step) {
        if constexpr (kHasTsit5 && !MIX) {
#pragma unroll
          for (int c = 0; c < N; c++) k1[c] = k64[6][c];
        } else if constexpr (kHasTsit5 && MIX) {
#pragma unroll
          for (int c = 0; c < N; c++) k1f[c] = k32[6][c];
        }
      }
      // savevalues! of the step end (not forced): the saveat points are
      // stored above; save_everystep adds (t, u) unless the last saved point
      // has this time.  Returns savedexactly: this call saved a point at t
      // (a saveat point counts in the first call only).
      [[maybe_unused]] auto step_point = [&]() -> bool {
        bool se = exact;
        exact = false;
        if constexpr (kEveryStep) {
          if (t != tlast) {
            save_event(a, u);
            tlast = t;
            se = true;
          }
        }
        return se;
      };
      if constexpr (kNcb > 0) {
        // apply_callback!: the continuous event's affect at its time, then
        // every discrete callback whose condition holds at the step end;
        // any affect re-evaluates the FSAL stage (the state may have moved).
        // Each applied callback first runs savevalues! of the step end, then
        // save_positions: (t, u) just before the affect unless that call
        // saved a point at t, and / or just after it.  A step with no
        // callback applied runs savevalues! once after them.
        bool fired = false, term = false;
        if (ev >= 0) {
          const bool se = step_point();
          if ((M::cb_save(ev) & 1) && !se) save_event(a, u);
          M::cb_affect(ev, u, p, t);
          if (M::cb_save(ev) & 2) save_event(a, u);
          if (M::cb_save(ev)) tlast = t;
          fired = true;
          term = term || M::cb_term(ev);
          nevent++;
        }
        for (int k = 0; k < kNcb; k++) {
          if (M::cb_kind(k) == 1 && M::cb_g(k, u, p, t) != 0.0) {
            const bool se = step_point();
            if ((M::cb_save(k) & 1) && !se) save_event(a, u);
            M::cb_affect(k, u, p, t);
            if (M::cb_save(k) & 2) save_event(a, u);
            if (M::cb_save(k)) tlast = t;
            fired = true;
            term = term || M::cb_term(k);
            nevent++;
          }
        }
        if (!fired) step_point();
        if (fired) {
          tevent = t;
          fsal_init();
          nf += 1;
        }
        for (int k = 0; k < kNcb; k++)
          if (M::cb_kind(k) == 0) gprev[k] = M::cb_g(k, u, p, t);
        if (term) {
          rc = RC_TERMINATED;
          return;
        }
      } else if constexpr (kEveryStep) {
        step_point();
      }
      bool bad = false;
#pragma unroll
      for (int c = 0; c < N; c++) bad = bad || (u1[c] != u1[c]);
      if (bad) {
        rc = RC_UNSTABLE;
        return;
      }
      dt = fmin(dtmax, dtnew);
    } else {
      nreject++;
      if constexpr (!FAST)
        dt = dt / fmin(1.0 / kQmin, q11 / kGammaSafety);
      else
        dt = dt * fast_exp2(-fmin(kL2QminInv, l11 - kL2Gamma));
    }
    if (!(tdir * (tf - t) > 0.0)) rc = RC_SUCCESS;
  }

  // ------------------------------------------------------------ finish
  __device__ void finish(const LaneArgs& a) {
    if constexpr (kHandoff) {
      handoff_append(a, i, rc == RC_HANDOFF);
      if (rc == RC_HANDOFF) return;  // the second pass writes every output
    }
    const int64_t NT = a.ntraj;
    if (a.us) {
      double nanv[N];
#pragma unroll
      for (int c = 0; c < N; c++) nanv[c] = __builtin_nan("");
      if (!back)
        for (int k = isave; k < a.ns; k++) store_save(a, k, nanv);
      else
        for (int k = isave; k >= 0; k--) store_save(a, k, nanv);
    }
    if (a.u_final)
#pragma unroll
      for (int c = 0; c < N; c++) a.u_final[(int64_t)c * NT + i] = u[c];
    if (a.t_final) a.t_final[i] = t;
    if (a.retcode) a.retcode[i] = rc;
    if (a.naccept) a.naccept[i] = naccept;
    if (a.nreject) a.nreject[i] = nreject;
    if (a.nf) a.nf[i] = nf;
    if (a.nstiff) a.nstiff[i] = nstiff;
    if (a.dt_first) a.dt_first[i] = dtfirst;
    if (a.nevent) a.nevent[i] = nevent;
    if (a.t_event) a.t_event[i] = tevent;
    if constexpr (kExtra) {  // event / step points: models that make them
      if (a.ev_n) a.ev_n[i] = nevsave;
      if (a.ev_t) {
        // slots past this trajectory's event points hold NaN
        for (int k = nevsave; k < a.ev_cap; k++) {
          const int64_t at = (int64_t)i * a.ev_cap + k;
          a.ev_t[at] = __builtin_nan("");
#pragma unroll
          for (int c = 0; c < N; c++) a.ev_u[at * N + c] = __builtin_nan("");
        }
      }
    }
    if constexpr (kNout > 0) {
      if (a.out) {
        double o[kNout];
        M::output(o, u, p, t, (int)rc);
#pragma unroll
        for (int j = 0; j < kNout; j++) a.out[(int64_t)j * NT + i] = o[j];
      }
    }
  }
};

// Kernel bodies, shared by the built-in kernels and the run-time compiled
// ones.  lane: one thread per trajectory of the launch order.
template <class M, bool MIX, int ALG, bool FAST>
__device__ __forceinline__ void lane_kernel_run(const LaneArgs& a) {
  const int64_t nq = lane_queue_len(a);
  if ((int64_t)blockIdx.x * blockDim.x >= nq) return;  // whole block idle
  lane_stage_saveat(a);
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  int64_t i = lane_traj(a, k);
  if (i < 0) return;
  Lane<M, MIX, ALG, FAST> L;
  L.init(a, i);
  while (L.rc == RC_DEFAULT) L.step(a);
  L.finish(a);
}

// persistent: lanes refill from the atomic queue, one atomic per wave.
template <class M, bool MIX, int ALG, bool FAST>
__device__ __forceinline__ void persistent_kernel_run(const LaneArgs& a) {
  constexpr int kRefill = 4;  // measured on 1M Lorenz: 1 -> 16.6 ms, 4 -> 14.9, 8 -> 15.1, 16 -> 15.4
  const int lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  lane_stage_saveat(a);
  const unsigned long long nq = (unsigned long long)lane_queue_len(a);
  Lane<M, MIX, ALG, FAST> L;
  bool active = false;
  bool drained = false;
  while (true) {
    if (!drained) {
      unsigned long long need = __ballot(!active);
      // refill in batches of kRefill idle lanes (or when every lane is idle):
      // a refill runs init (initdt, first saves, FSAL) for the new lanes with
      // the rest of the wave masked off, plus one queue atomic
      if (__popcll(need) >= kRefill || need == ~0ull) {
        int cnt = __popcll(need);
        int leader = __builtin_ctzll(need);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.queue, (unsigned long long)cnt);
        base = __shfl(base, leader);
        if (base + (unsigned long long)cnt >= nq) drained = true;
        if (!active) {
          unsigned long long my = base + (unsigned long long)__popcll(need & lt_mask);
          int64_t i = my < nq ? lane_traj(a, (int64_t)my) : -1;
          if (i >= 0) {
            L.init(a, i);
            if (L.rc == RC_DEFAULT) {
              active = true;
            } else {
              L.finish(a);
            }
          }
        }
      }
    }
    if (!__any(active)) {
      if (drained) break;
      continue;
    }
    if (active) {
      L.step(a);
      if (L.rc != RC_DEFAULT) {
        L.finish(a);
        active = false;
      }
    }
  }
}

}  // namespace mideq
