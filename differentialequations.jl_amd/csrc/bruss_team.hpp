// bruss_team.hpp -- uneven Brusselator batch: one persistent kernel, a
// team of L lanes per instance (C = 128/L state components per lane), work
// pulled from an atomic queue in a host-chosen order.
//
// Per instance (state size n = 2*Ngrid, interleaved u_j, v_j):
//   * RHS: each lane evaluates its components; neighbours come from the
//     team's LDS slot (periodic indices).
//   * Norms: per-lane partial sums of squares, then a butterfly tree across
//     the team (shfl_xor), the order the CPU restatement reproduces.
//   * Rodas5: W = J - I/(gamma dt) in the team's LDS in band + border form
//     (16 NK doubles), LU with partial pivoting over the structurally
//     nonzero entries only (each of them sees the same fma sequence as in a
//     dense sequential LU), triangular solves as register chains.
// Stage arithmetic is FP64 or, in the mixed form, FP32 (RHS and stage
// sums); state update, norm, time, J, W and LU are FP64.
#pragma once

#include <type_traits>

#include "bruss.hpp"
#include "device_common.hpp"
#include "lane_solver.hpp"

namespace mideq {

// Make this wave's earlier global/LDS writes visible to the other lanes of
// the wave and keep the compiler from moving memory operations across.
__device__ __forceinline__ void team_sync() {
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
}

template <int L>
__device__ __forceinline__ double team_sum(double a) {
#pragma unroll
  for (int w = L / 2; w >= 1; w >>= 1) a = a + __shfl_xor(a, w, L);
  return a;
}

template <int L>
__device__ __forceinline__ double team_max(double a) {
#pragma unroll
  for (int w = L / 2; w >= 1; w >>= 1) {
    double o = __shfl_xor(a, w, L);
    a = (o > a) ? o : a;
  }
  return a;
}

template <int L>
__device__ __forceinline__ int team_or(int v) {
#pragma unroll
  for (int w = L / 2; w >= 1; w >>= 1) v |= __shfl_xor(v, w, L);
  return v;
}

template <int L>
__device__ __forceinline__ int team_min_int(int v) {
#pragma unroll
  for (int w = L / 2; w >= 1; w >>= 1) v = min(v, __shfl_xor(v, w, L));
  return v;
}

// A value that is the same in every lane of the team, marked uniform to the
// compiler when the team is the whole wavefront (scalar registers, scalar
// branches); a 32-lane team shares its wavefront with another instance, so
// its values stay per lane.
template <int L>
__device__ __forceinline__ int team_uni(int v) {
  if constexpr (L == 64)
    return __builtin_amdgcn_readfirstlane(v);
  else
    return v;
}

template <int L>
__device__ __forceinline__ unsigned long long team_uni_u64(unsigned long long v) {
  if constexpr (L == 64) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
  } else {
    return v;
  }
}

template <int L>
__device__ __forceinline__ int team_and(int v) {
#pragma unroll
  for (int w = L / 2; w >= 1; w >>= 1) v &= __shfl_xor(v, w, L);
  return v;
}

// SAVE: builds with saveat (the step's stages stay live until the accepted
// step has written its saved points).  Every member function is inlined by
// force: an out-of-line call passes the team through scratch memory, where
// its LDS pointers turn into generic ones (flat loads instead of LDS).
template <int L, int C, bool MIX, int ALG, bool FAST, bool SAVE>
struct BTeam {
  static constexpr bool kAuto = (ALG == ALG_AUTO || ALG == ALG_AUTO5P);
  static constexpr bool kHasTsit5 = (ALG != ALG_RODAS5 && ALG != ALG_RODAS5P);
  static constexpr bool kHasRodas = (ALG != ALG_TSIT5);
  static constexpr bool kP = (ALG == ALG_RODAS5P || ALG == ALG_AUTO5P);  // Rodas5P tableau

  int tl;
  int64_t i;
  int n, N;
  double A, B, B1, D;
  float Af, Bf, B1f, Df;
  double u[C];
  double k1[C];
  float k1f[C];
  double t, tcomp, tf, dt, dtmax, qold, qoldpow, qsteady_max, eig, dtfirst;
  double lqold;  // fast controller: log2 of the previous accepted EEst (>= log2 1e-4)
  double tsnap;  // 100 eps(tf): a forward step ending this close to tf ends at tf
  double tdir;   // +1, or -1 for tf < t0: integrated backward with steps of -dt (dt > 0)
  int64_t iter;
  int32_t naccept, nreject, nf, nstiff, count, nanrej;
  int32_t isave;  // next saveat point (SAVE builds)
  double tnext;   // its time, +inf (-inf backward) once every point is saved
  int8_t rc, mode;
  double* ldsd;  // team LDS: exchange vector (BRUSS_NMAX doubles)
  double* ldsr;  // team LDS: reciprocal pivots
  float* ldsf;   // team LDS: FP32 exchange vector
  int* ldsi;     // team LDS: pivot rows
  int nb;        // n - kT: positions of W stored in band form
  double* Wb;    // team LDS (Rosenbrock kernels): W band rows
  double* Wr;    //   their entries in the last kT columns
  double* Wd;    //   the last kT rows, dense
  bool lu_swapped;  // the last factorization interchanged rows

  // Lane tl owns grid points j = tl + L*g (g < C/2), each as the pair of
  // components (u_j, v_j) = (2j, 2j + 1) in slots (2g, 2g + 1), so both rows
  // of a grid point's RHS are evaluated by one lane with no exchange.
  static_assert(C % 2 == 0, "slots hold (u, v) pairs");
  __device__ __forceinline__ int comp(int m) const { return 2 * (tl + L * (m >> 1)) + (m & 1); }
  // Rows of W (its construction and the border pass of the back
  // substitution) are owned lane-strided.
  __device__ __forceinline__ int rowc(int m) const { return tl + L * m; }

  // ---------------------------------------------------------------- RHS
  __device__ __forceinline__ void rhs64(double* out, const double* in) {
#pragma unroll
    for (int m = 0; m < C; m++) {
      int c = comp(m);
      if (c < n) ldsd[c] = in[m];
    }
    team_sync();
#pragma unroll
    for (int g = 0; g < C / 2; g++) {
      int cu = comp(2 * g);  // u of grid point j = cu / 2; v is cu + 1 (n is even)
      double du = 0.0, dv = 0.0;
      if (cu < n) {
        int j = cu >> 1;
        int jm = (j == 0) ? N - 1 : j - 1;
        int jp = (j == N - 1) ? 0 : j + 1;
        double uu = in[2 * g], vv = in[2 * g + 1];
        double uuv = uu * uu * vv;
        double lu = ldsd[2 * jm] - 2.0 * uu + ldsd[2 * jp];
        double lv = ldsd[2 * jm + 1] - 2.0 * vv + ldsd[2 * jp + 1];
        du = A + uuv - B1 * uu + D * lu;
        dv = B * uu - uuv + D * lv;
      }
      out[2 * g] = du;
      out[2 * g + 1] = dv;
    }
    team_sync();
  }

  __device__ __forceinline__ void rhs32(float* out, const float* in) {
#pragma unroll
    for (int m = 0; m < C; m++) {
      int c = comp(m);
      if (c < n) ldsf[c] = in[m];
    }
    team_sync();
#pragma unroll
    for (int g = 0; g < C / 2; g++) {
      int cu = comp(2 * g);
      float du = 0.0f, dv = 0.0f;
      if (cu < n) {
        int j = cu >> 1;
        int jm = (j == 0) ? N - 1 : j - 1;
        int jp = (j == N - 1) ? 0 : j + 1;
        float uu = in[2 * g], vv = in[2 * g + 1];
        float uuv = uu * uu * vv;
        float lu = ldsf[2 * jm] - 2.0f * uu + ldsf[2 * jp];
        float lv = ldsf[2 * jm + 1] - 2.0f * vv + ldsf[2 * jp + 1];
        du = Af + uuv - B1f * uu + Df * lu;
        dv = Bf * uu - uuv + Df * lv;
      }
      out[2 * g] = du;
      out[2 * g + 1] = dv;
    }
    team_sync();
  }

  __device__ __forceinline__ void fsal_init() {
    rhs64(k1, u);
    if (MIX) {
      float uf[C];
#pragma unroll
      for (int m = 0; m < C; m++) uf[m] = (float)u[m];
      rhs32(k1f, uf);
    }
  }

  // RMS of owned values x[m]/sc[m] style partial sums -> sqrt(sum/n)
  __device__ __forceinline__ double rms_from_partial(double part) {
    return sqrt(team_sum<L>(part) / (double)n);
  }

  // ---------------------------------------------------------------- initdt
  __device__ __forceinline__ double initdt(const BrussArgs& a, double t0) {
    double dtmin_e = nextafter(fmax(a.dtmin, jl_eps(t0)), __builtin_inf());
    double smalldt = fmax(dtmin_e, 1e-6);
    double sk[C];
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int m = 0; m < C; m++) {
      sk[m] = 1.0;
      if (comp(m) < n) {
        sk[m] = fma(fabs(u[m]), a.reltol, a.abstol);
        double x = u[m] / sk[m];
        s0 = s0 + x * x;
        double y = k1[m] / sk[m];
        s1 = s1 + y * y;
      }
    }
    double d0 = rms_from_partial(s0);
    double d1 = rms_from_partial(s1);
    if (d1 != d1) return dtmin_e;
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : (d0 / d1) / 100.0;
    dt0 = fmin(dt0, dtmax);
    if (dt0 < 10.0 * 2.220446049250313e-16) return smalldt;
    double u1[C], f1[C];
    const double sdt0 = tdir * dt0;  // the trial step in the direction of integration
#pragma unroll
    for (int m = 0; m < C; m++) u1[m] = fma(sdt0, k1[m], u[m]);
    rhs64(f1, u1);
    nf += 1;
    int same = 1;
#pragma unroll
    for (int m = 0; m < C; m++)
      if (comp(m) < n) same &= (k1[m] == f1[m]) ? 1 : 0;
    if (team_and<L>(same)) return fmax(dtmin_e, 100.0 * dt0);
    double s2 = 0.0;
#pragma unroll
    for (int m = 0; m < C; m++) {
      if (comp(m) < n) {
        double x = (f1[m] - k1[m]) / sk[m];
        s2 = s2 + x * x;
      }
    }
    double d2 = rms_from_partial(s2) / dt0;
    double mx = fmax(d1, d2);
    double dt1;
    if (mx <= 1e-15)
      dt1 = fmax(1e-6, dt0 * 1e-3);
    else
      dt1 = ctl_exp2(-(2.0 * 3.32192809488736234787 + ctl_log2(mx)) / 5.0);
    return fmax(dtmin_e, fmin(fmin(100.0 * dt0, dt1), dtmax));
  }

  // ---------------------------------------------------------------- init
  __device__ __forceinline__ void init(const BrussArgs& a, int64_t idx) {
    i = idx;
    const int64_t NT = a.ntraj;
    n = team_uni<L>(a.nvec ? a.nvec[i] : a.nmax);
    N = n / 2;
    nb = n - kT;
    A = a.p[i];
    B = a.p[NT + i];
    double al = a.p[2 * NT + i];
    D = al * (double)(N * N);
    B1 = B + 1.0;
    Af = (float)A;
    Bf = (float)B;
    float alf = (float)al;
    Df = alf * (float)(N * N);
    B1f = Bf + 1.0f;
    double t0 = a.t0 ? a.t0[i] : a.tspan0;
    tf = a.tf ? a.tf[i] : a.tspan1;
    tdir = tf < t0 ? -1.0 : 1.0;
    t = t0;
    tcomp = 0.0;
    dtmax = a.dtmax > 0 ? a.dtmax : fabs(tf - t0);
    naccept = nreject = nstiff = 0;
    nf = 1;
    iter = 0;
    count = 0;
    nanrej = 0;
    eig = 0.0;
    dtfirst = 0.0;
    rc = RC_DEFAULT;
    isave = 0;
    tnext = __builtin_inf();
    mode = (ALG == ALG_RODAS5 || ALG == ALG_RODAS5P) ? 1 : 0;
    if (kAuto && a.alg_hint) mode = a.alg_hint[i] ? 1 : 0;
    qsteady_max = (a.qsteady_default && mode == 1) ? 1.2 : 1.0;
    // n > L * C: more components than the team holds (the host routes only
    // instances that fit; this keeps a wrong routing in bounds)
    if (n < BRUSS_NMIN || n > a.nmax || n > L * C || (n & 1)) {
      rc = RC_INITIALFAILURE;
      n = 0;
      return;
    }
#pragma unroll
    for (int m = 0; m < C; m++) {
      int c = comp(m);
      u[m] = (c < n) ? a.u0[(int64_t)c * NT + i] : 0.0;
    }
    fsal_init();
    if (a.dt_fixed > 0)
      dt = a.dt_fixed;
    else if (a.dt0)
      dt = a.dt0[i];
    else
      dt = initdt(a, t0);
    qold = kQoldInit;
    qoldpow = ctl_pow(kQoldInit, kBeta2);
    lqold = kL2QoldInit;
    // a step ending within 100 eps(max(t, tf)) of tf ends at tf, t the
    // step's start: forward t < tf in the loop, so the window is fixed;
    // backward max(t, tf) = t and the window is taken per step
    tsnap = 100.0 * jl_eps(tf);
    // a zero-length span is done; a NaN end is not integrated
    if (tf == t0)
      rc = RC_SUCCESS;
    else if (!(tf > t0) && !(tf < t0))
      rc = RC_INITIALFAILURE;
    if constexpr (SAVE) {
      // saveat points (ascending) before t0 in the direction of integration
      // are unreachable; a point at t0 is u0 itself.  Backward, the points
      // are visited from the last one down.
      double nanv[C];
#pragma unroll
      for (int m = 0; m < C; m++) nanv[m] = __builtin_nan("");
      if (tdir > 0.0) {
        while (isave < a.ns && a.saveat[isave] < t0) store_save(a, isave++, nanv);
        while (isave < a.ns && a.saveat[isave] == t0) store_save(a, isave++, u);
      } else {
        isave = a.ns - 1;
        while (isave >= 0 && a.saveat[isave] > t0) store_save(a, isave--, nanv);
        while (isave >= 0 && a.saveat[isave] == t0) store_save(a, isave--, u);
      }
      set_tnext(a);
    }
  }

  __device__ __forceinline__ void set_tnext(const BrussArgs& a) {
    if (tdir > 0.0)
      tnext = isave < a.ns ? a.saveat[isave] : __builtin_inf();
    else
      tnext = isave >= 0 ? a.saveat[isave] : -__builtin_inf();
  }

  // ---------------------------------------------------------------- saveat
  // Point k of this instance: owned components c < n at us[(k*nmax + c)*ntraj
  // + i] (layout 0) or us[(i*ns + k)*nmax + c] (layout 1, component fastest:
  // a team's store is contiguous), FP64 or FP32.
  __device__ __forceinline__ void store_save(const BrussArgs& a, int k, const double* v) {
#pragma unroll
    for (int m = 0; m < C; m++) {
      const int c = comp(m);
      if (c < n) {
        const int64_t at = a.save_layout ? ((int64_t)i * a.ns + k) * a.nmax + c
                                         : ((int64_t)k * a.nmax + c) * a.ntraj + i;
        if (a.save_f32)
          ((float*)a.us)[at] = (float)v[m];
        else
          ((double*)a.us)[at] = v[m];
      }
    }
  }

  // ---------------------------------------------------------------- Tsit5
  // EEst (reference controller) or EEst^2 = mean(r^2) without the square
  // root (fast controller, which works on log2 EEst^2).
  __device__ __forceinline__ double err_norm(const double* ut, const double* u1, double abstol,
                             double reltol) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < C; m++) {
      if (comp(m) < n) {
        double sc = fma(fmax(fabs(u[m]), fabs(u1[m])), reltol, abstol);
        double r = ut[m] / sc;
        s = s + r * r;
      }
    }
    if constexpr (FAST) return team_sum<L>(s) / (double)n;
    return rms_from_partial(s);
  }

  __device__ __forceinline__ void eig_estimate(const double* k7, const double* k6, const double* u1,
                               const double* g6) {
    double e = 0.0;
    int bad = 0;
#pragma unroll
    for (int m = 0; m < C; m++) {
      if (comp(m) < n) {
        double r = fabs((k7[m] - k6[m]) / (u1[m] - g6[m]));
        bad |= (r != r) ? 1 : 0;
        e = (r > e) ? r : e;
      }
    }
    e = team_max<L>(e);
    eig = team_or<L>(bad) ? __builtin_nan("") : e;
  }

  __device__ __forceinline__ double tsit5_step64(double dts, double* u1, double abstol, double reltol,
                                double (*kout)[C]) {
    using namespace ts;
    double k[7][C];
    double g[C], g6[C], ut[C];
    double aa = dts * a21;
#pragma unroll
    for (int m = 0; m < C; m++) {
      k[0][m] = k1[m];
      g[m] = fma(aa, k[0][m], u[m]);
    }
    rhs64(k[1], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = a32 * k[1][m];
      acc = fma(a31, k[0][m], acc);
      g[m] = fma(dts, acc, u[m]);
    }
    rhs64(k[2], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = a43 * k[2][m];
      acc = fma(a42, k[1][m], acc);
      acc = fma(a41, k[0][m], acc);
      g[m] = fma(dts, acc, u[m]);
    }
    rhs64(k[3], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = a54 * k[3][m];
      acc = fma(a53, k[2][m], acc);
      acc = fma(a52, k[1][m], acc);
      acc = fma(a51, k[0][m], acc);
      g[m] = fma(dts, acc, u[m]);
    }
    rhs64(k[4], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = a65 * k[4][m];
      acc = fma(a64, k[3][m], acc);
      acc = fma(a63, k[2][m], acc);
      acc = fma(a62, k[1][m], acc);
      acc = fma(a61, k[0][m], acc);
      g6[m] = fma(dts, acc, u[m]);
    }
    rhs64(k[5], g6);
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = a76 * k[5][m];
      acc = fma(a75, k[4][m], acc);
      acc = fma(a74, k[3][m], acc);
      acc = fma(a73, k[2][m], acc);
      acc = fma(a72, k[1][m], acc);
      acc = fma(a71, k[0][m], acc);
      u1[m] = fma(dts, acc, u[m]);
    }
    rhs64(k[6], u1);
    if (kAuto) eig_estimate(k[6], k[5], u1, g6);
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = bt7 * k[6][m];
      acc = fma(bt6, k[5][m], acc);
      acc = fma(bt5, k[4][m], acc);
      acc = fma(bt4, k[3][m], acc);
      acc = fma(bt3, k[2][m], acc);
      acc = fma(bt2, k[1][m], acc);
      acc = fma(bt1, k[0][m], acc);
      ut[m] = dts * acc;
    }
#pragma unroll
    for (int m = 0; m < C; m++) k1[m] = k[6][m];  // FSAL candidate (restored on reject)
    if constexpr (SAVE) {
#pragma unroll
      for (int s = 0; s < 7; s++)
#pragma unroll
        for (int m = 0; m < C; m++) kout[s][m] = k[s][m];
    }
    return err_norm(ut, u1, abstol, reltol);
  }

  __device__ __forceinline__ double tsit5_step_mix(double dts, double* u1, float* k7out, double abstol,
                                   double reltol, float (*kout)[C]) {
    using namespace ts;
    float k[7][C];
    float u32[C], g[C], g6[C], u1f[C];
    double ut[C];
    float d32 = (float)dts;
#pragma unroll
    for (int m = 0; m < C; m++) {
      u32[m] = (float)u[m];
      k[0][m] = k1f[m];
    }
    float aa = d32 * (float)a21;
#pragma unroll
    for (int m = 0; m < C; m++) g[m] = fmaf(aa, k[0][m], u32[m]);
    rhs32(k[1], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      float acc = (float)a32 * k[1][m];
      acc = fmaf((float)a31, k[0][m], acc);
      g[m] = fmaf(d32, acc, u32[m]);
    }
    rhs32(k[2], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      float acc = (float)a43 * k[2][m];
      acc = fmaf((float)a42, k[1][m], acc);
      acc = fmaf((float)a41, k[0][m], acc);
      g[m] = fmaf(d32, acc, u32[m]);
    }
    rhs32(k[3], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      float acc = (float)a54 * k[3][m];
      acc = fmaf((float)a53, k[2][m], acc);
      acc = fmaf((float)a52, k[1][m], acc);
      acc = fmaf((float)a51, k[0][m], acc);
      g[m] = fmaf(d32, acc, u32[m]);
    }
    rhs32(k[4], g);
#pragma unroll
    for (int m = 0; m < C; m++) {
      float acc = (float)a65 * k[4][m];
      acc = fmaf((float)a64, k[3][m], acc);
      acc = fmaf((float)a63, k[2][m], acc);
      acc = fmaf((float)a62, k[1][m], acc);
      acc = fmaf((float)a61, k[0][m], acc);
      g6[m] = fmaf(d32, acc, u32[m]);
    }
    rhs32(k[5], g6);
#pragma unroll
    for (int m = 0; m < C; m++) {
      float acc = (float)a76 * k[5][m];
      acc = fmaf((float)a75, k[4][m], acc);
      acc = fmaf((float)a74, k[3][m], acc);
      acc = fmaf((float)a73, k[2][m], acc);
      acc = fmaf((float)a72, k[1][m], acc);
      acc = fmaf((float)a71, k[0][m], acc);
      u1[m] = fma(dts, (double)acc, u[m]);
      u1f[m] = (float)u1[m];
    }
    rhs32(k[6], u1f);
    if (kAuto) {
      double k7d[C], k6d[C], g6d[C];
#pragma unroll
      for (int m = 0; m < C; m++) {
        k7d[m] = (double)k[6][m];
        k6d[m] = (double)k[5][m];
        g6d[m] = (double)g6[m];
      }
      eig_estimate(k7d, k6d, u1, g6d);
    }
#pragma unroll
    for (int m = 0; m < C; m++) {
      float acc = (float)bt7 * k[6][m];
      acc = fmaf((float)bt6, k[5][m], acc);
      acc = fmaf((float)bt5, k[4][m], acc);
      acc = fmaf((float)bt4, k[3][m], acc);
      acc = fmaf((float)bt3, k[2][m], acc);
      acc = fmaf((float)bt2, k[1][m], acc);
      acc = fmaf((float)bt1, k[0][m], acc);
      ut[m] = dts * (double)acc;
      k7out[m] = k[6][m];
    }
    if constexpr (SAVE) {
#pragma unroll
      for (int s = 0; s < 7; s++)
#pragma unroll
        for (int m = 0; m < C; m++) kout[s][m] = k[s][m];
    }
    return err_norm(ut, u1, abstol, reltol);
  }

  // ---------------------------------------------------------------- Rodas5
  // W = J - I/(gamma dt) lives in the team's LDS in band + border form.  The
  // Brusselator's J (interleaved u, v; periodic) has its nonzeros in
  // columns r-2..r+2 of row r plus the wrap entries of rows 0, 1, n-2, n-1.
  // During LU with partial pivoting every row still to be factored at step k
  // keeps its nonzeros in columns [k, k+4] and the last kT columns, and the
  // pivot comes from positions k, k+1, k+2 or the last kT rows.  Storage by
  // position p (nb = n - kT):
  //   p <  nb: Wb[p*8 + (c - p + 2)] for c in [p-2, p+4], c < nb (the two
  //            entries left of the diagonal end up holding L multipliers)
  //            Wr[p*kT + (c - nb)]  for c >= nb
  //   p >= nb: Wd[(p - nb)*NK + c]  every column (dense rows)
  // The factorization and the solves perform exactly the operations of the
  // dense LU that involve a stored entry, in the same order per entry; the
  // others multiply an exact zero and leave values unchanged, unless a
  // non-finite value meets that zero (then the step is marked failed,
  // EEst = NaN, as the dense factorization's NaNs would make it).  Row
  // interchanges move columns >= k only (LINPACK form): the forward
  // elimination applies each interchange when it reaches it, which per entry
  // is the same sequence of operations as permuting first and then
  // eliminating with the fully interchanged L.
  static constexpr int kT = 4;
  static constexpr int NK = L * C;

  __device__ __forceinline__ double* wp(int p, int c) const {
    if (p >= nb) return Wd + (p - nb) * NK + c;
    if (c >= nb) return Wr + p * kT + (c - nb);
    return Wb + p * 8 + (c - p + 2);
  }
  // pivot candidates at step k: k, then k+1, k+2 while < nb, then the
  // dense positions past k, ascending (nbc band candidates past k)
  __device__ __forceinline__ int nband_after(int k) const { return max(0, min(2, nb - 1 - k)); }
  __device__ __forceinline__ int cand(int k, int nbc, int i) const {
    return i == 0 ? k : (i <= nbc ? k + i : max(nb, k + 1) + (i - 1 - nbc));
  }
  // columns of the pivot row's pattern: k, then k+1.. while < nb and <= k+4
  // (nwc of them), then the border columns past k
  __device__ __forceinline__ int nwin_after(int k) const { return max(0, min(4, nb - 1 - k)); }
  __device__ __forceinline__ int patcol(int k, int nwc, int j) const {
    return j == 0 ? k : (j <= nwc ? k + j : max(nb, k + 1) + (j - 1 - nwc));
  }

  // Builds W and factors it.  0: factored; 1: a zero pivot came first
  // (EEst = 2, as the dense LU's early return); 2: a non-finite value met a
  // skipped zero first (EEst = NaN).
  __device__ __forceinline__ int factor_band(double dts) {
    // the sizes are the team's (uniform: scalar loop control on 64-lane teams)
    n = team_uni<L>(n);
    nb = team_uni<L>(nb);
#pragma unroll
    for (int m = 0; m < C; m++) {
      int c = comp(m);
      if (c < n) ldsd[c] = u[m];
    }
    for (int e = tl; e < 8 * nb; e += L) Wb[e] = 0.0;
    for (int e = tl; e < kT * nb; e += L) Wr[e] = 0.0;
    for (int e = tl; e < kT * NK; e += L) Wd[e] = 0.0;
    // pivot rows: a column without an interchange keeps this entry
    for (int e = tl; e < n; e += L) ldsi[e] = e;
    team_sync();
    const double dtgamma = dts * rd_gamma<kP>();
    const double w = 1.0 / dtgamma;
    double best = 0.0;
#pragma unroll
    for (int m = 0; m < C; m++) {
      int r = rowc(m);
      if (r < n) {
        int j = r >> 1;
        int jm = (j == 0) ? N - 1 : j - 1;
        int jp = (j == N - 1) ? 0 : j + 1;
        double uu = ldsd[2 * j], vv = ldsd[2 * j + 1];
        double uv2 = 2.0 * uu * vv, u2 = uu * uu;
        // the row's four J entries (0.0 + value, as J accumulates into a
        // zeroed matrix) in ascending column order; e_d the diagonal one
        int c0, c1, c2, c3, id;
        double e0, e1, e2, e3;
        const double dd = 0.0 + D;
        if ((r & 1) == 0) {
          const double a = 0.0 + (uv2 - B1 - 2.0 * D), b = 0.0 + u2;
          if (j == 0) {  // columns 0, 1, 2, n-2
            c0 = r; e0 = a; c1 = r + 1; e1 = b; c2 = 2 * jp; e2 = dd; c3 = 2 * jm; e3 = dd; id = 0;
          } else if (j == N - 1) {  // columns 0, n-4, n-2, n-1
            c0 = 2 * jp; e0 = dd; c1 = 2 * jm; e1 = dd; c2 = r; e2 = a; c3 = r + 1; e3 = b; id = 2;
          } else {
            c0 = 2 * jm; e0 = dd; c1 = r; e1 = a; c2 = r + 1; e2 = b; c3 = 2 * jp; e3 = dd; id = 1;
          }
        } else {
          const double a = 0.0 + (B - uv2), b = 0.0 + (-u2 - 2.0 * D);
          if (j == 0) {  // columns 0, 1, 3, n-1
            c0 = r - 1; e0 = a; c1 = r; e1 = b; c2 = 2 * jp + 1; e2 = dd; c3 = 2 * jm + 1; e3 = dd; id = 1;
          } else if (j == N - 1) {  // columns 1, n-3, n-2, n-1
            c0 = 2 * jp + 1; e0 = dd; c1 = 2 * jm + 1; e1 = dd; c2 = r - 1; e2 = a; c3 = r; e3 = b; id = 3;
          } else {
            c0 = 2 * jm + 1; e0 = dd; c1 = r - 1; e1 = a; c2 = r; e2 = b; c3 = 2 * jp + 1; e3 = dd; id = 2;
          }
        }
        if (kAuto) {
          double sr = 0.0 + fabs(e0);
          sr = sr + fabs(e1);
          sr = sr + fabs(e2);
          sr = sr + fabs(e3);
          best = (sr > best) ? sr : best;
        }
        *wp(r, c0) = (id == 0) ? e0 - w : e0;
        *wp(r, c1) = (id == 1) ? e1 - w : e1;
        *wp(r, c2) = (id == 2) ? e2 - w : e2;
        *wp(r, c3) = (id == 3) ? e3 - w : e3;
      }
    }
    if (kAuto) eig = team_max<L>(best);
    team_sync();
    lu_swapped = false;
    // columns k <= nb - 5: the candidates and the pattern have their fixed
    // shape (rows k, k+1, k+2 and the kT dense rows; columns k..k+4 and the
    // border), so column k's seven candidate values load at once
    const int kmain = nb - 5;
    // this lane's entries of the trailing update (6 candidate rows x 8
    // pattern columns) and of the multiplier column, as offsets from Wb
    // affine in k (the band, border and dense parts are contiguous)
    constexpr int kE = (48 + L - 1) / L;
    int a_off[kE], a_st[kE], l_off[kE], l_st[kE], w_off[kE], w_st[kE];
#pragma unroll
    for (int e = 0; e < kE; e++) {
      const int pi = tl + e * L;
      const int qi = (pi >> 3) < 6 ? (pi >> 3) : 5, cj = 1 + (pi & 7);
      lu_pattern_offsets(qi, cj, a_off[e], a_st[e], l_off[e], l_st[e], w_off[e], w_st[e]);
    }
    int m_off = 0, m_st = 0;
    {
      int t0, t1, t2, t3;
      lu_pattern_offsets(tl < 6 ? tl : 5, 1, t0, t1, m_off, m_st, t2, t3);
    }
    for (int k = 0; k <= kmain; k++) {
      const double v0 = Wb[k * 8 + 2], v1 = Wb[(k + 1) * 8 + 1], v2 = Wb[(k + 2) * 8];
      const double v3 = Wd[k], v4 = Wd[NK + k], v5 = Wd[2 * NK + k], v6 = Wd[3 * NK + k];
      // this lane's operands of the trailing update and of the multiplier
      // column without an interchange, loaded with the candidates: their
      // LDS latency overlaps the pivot choice and 1/d (nothing writes W
      // before the update below; an interchange reloads them)
      double lq[kE], wq[kE], aq[kE];
#pragma unroll
      for (int e = 0; e < kE; e++) {
        if (tl + e * L < 48) {
          lq[e] = Wb[l_off[e] + k * l_st[e]];
          wq[e] = Wb[w_off[e] + k * w_st[e]];
          aq[e] = Wb[a_off[e] + k * a_st[e]];
        }
      }
      const double mq = tl < 6 ? Wb[m_off + k * m_st] : 0.0;
      // the first maximal |W(q, k)| in ascending row order
      double d = v0;
      int pk = k;
      const double vq[6] = {v1, v2, v3, v4, v5, v6};
#pragma unroll
      for (int t = 0; t < 6; t++) {
        const bool take = fabs(vq[t]) > fabs(d);
        d = take ? vq[t] : d;
        pk = take ? (t < 2 ? k + 1 + t : nb + (t - 2)) : pk;
      }
      pk = team_uni<L>(pk);
      if (pk != k) {
        lu_swapped = true;
        lu_column(k, 2, 4, 7, 9, pk, d);
        continue;
      }
      // no interchange (the usual case): the pattern's fixed addresses
      const double rr = 1.0 / d;
      if (tl == 0) ldsr[k] = rr;
#pragma unroll
      for (int e = 0; e < kE; e++) {
        if (tl + e * L < 48) {
          const double l = lq[e] * rr;
          Wb[a_off[e] + k * a_st[e]] = fma(-l, wq[e], aq[e]);
        }
      }
      if (tl < 6) Wb[m_off + k * m_st] = mq * rr;
      team_sync();
    }
    for (int k = kmain + 1; k < n; k++) {
      const int tail0 = max(nb, k + 1);
      const int nbc = nband_after(k);
      const int nc = 1 + nbc + (n - tail0);
      const int nwc = nwin_after(k);
      const int np = 1 + nwc + (n - tail0);
      // pivot: the first maximal |W(q, k)| over the candidates (every other
      // row of column k is an exact zero)
      double d = *wp(k, k);
      double bv = fabs(d);
      int pk = k;
      for (int i = 1; i < nc; i++) {
        const int q = cand(k, nbc, i);
        const double v = *wp(q, k);
        if (fabs(v) > bv) {
          bv = fabs(v);
          pk = q;
          d = v;
        }
      }
      pk = team_uni<L>(pk);
      lu_swapped = lu_swapped || pk != k;
      lu_column(k, nbc, nwc, nc, np, pk, d);
    }
    // The outcomes of the dense LU, from the final factors, lane-parallel
    // over the columns: its first zero pivot, and the first column where a
    // non-finite value meets one of the zeros skipped here (the dense LU
    // multiplies every row past k by 1/d and every column past k by the
    // pivot row).  Column k's inputs -- its pivot W(k, k), 1/d, the pivot
    // row over the pattern and the multipliers -- are final once column k is
    // done: later columns update and interchange rows and columns past k only.
    int sing_k = n, pois_k = n;
    for (int k = tl; k < n; k += L) {
      const int tail0 = max(nb, k + 1);
      const int nbc = nband_after(k);
      const int nc = 1 + nbc + (n - tail0);
      const int nwc = nwin_after(k);
      const int np = 1 + nwc + (n - tail0);
      if (*wp(k, k) == 0.0 && sing_k == n) sing_k = k;
      bool bad = false;
      if ((n - 1 - k) > (nc - 1)) {  // rows past k outside the candidates
        bad = !(fabs(ldsr[k]) <= __DBL_MAX__);
        for (int j = 1; j < np; j++) bad = bad || !(fabs(*wp(k, patcol(k, nwc, j))) <= __DBL_MAX__);
      }
      if ((n - 1 - k) > (np - 1))  // columns past k outside the pattern
        for (int i = 1; i < nc; i++) bad = bad || !(fabs(*wp(cand(k, nbc, i), k)) <= __DBL_MAX__);
      if (bad && pois_k == n) pois_k = k;
    }
    sing_k = team_min_int<L>(sing_k);
    pois_k = team_min_int<L>(pois_k);
    if (sing_k < n && sing_k <= pois_k) return 1;
    return pois_k < n ? 2 : 0;
  }

  // Offsets from Wb (affine in the column k: offset + k * stride) of the
  // fixed-pattern LU entries for candidate row index qi (0, 1: rows k+1,
  // k+2; 2..5: dense rows nb..nb+3) and pattern column index cj (1..4:
  // columns k+1..k+4; 5..8: border columns nb..nb+3): the updated entry
  // W(q, c), its multiplier source W(q, k) and the pivot row's W(k, c).
  __device__ __forceinline__ void lu_pattern_offsets(int qi, int cj, int& a_off, int& a_st,
                                                     int& l_off, int& l_st, int& w_off,
                                                     int& w_st) const {
    const bool qband = qi < 2, cband = cj <= 4;
    const int dq = qi + 1, r = qi - 2, dc = cj, cb = cj - 5;
    constexpr int oR = 8 * NK, oD = 8 * NK + kT * NK;
    // W(q, c)
    if (qband && cband) {
      a_off = 7 * dq + dc + 2;
      a_st = 8;
    } else if (qband) {
      a_off = oR + kT * dq + cb;
      a_st = kT;
    } else if (cband) {
      a_off = oD + r * NK + dc;
      a_st = 1;
    } else {
      a_off = oD + r * NK + nb + cb;
      a_st = 0;
    }
    // W(q, k)
    if (qband) {
      l_off = 7 * dq + 2;
      l_st = 8;
    } else {
      l_off = oD + r * NK;
      l_st = 1;
    }
    // W(k, c)
    if (cband) {
      w_off = dc + 2;
      w_st = 8;
    } else {
      w_off = oR + cb;
      w_st = kT;
    }
  }

  // Column k of the factorization once its pivot row pk (pivot value d) is
  // chosen: interchange rows k and pk over the pattern, store 1/d and pk,
  // update the trailing candidate rows x pattern columns, store the
  // multipliers.
  __device__ __forceinline__ void lu_column(int k, int nbc, int nwc, int nc, int np, int pk,
                                            double d) {
    if (pk != k) {
      for (int jc = tl; jc < np; jc += L) {
        const int c = patcol(k, nwc, jc);
        double* a = wp(k, c);
        double* b = wp(pk, c);
        const double x = *a, y = *b;
        *a = y;
        *b = x;
      }
      team_sync();
    }
    const double rr = 1.0 / d;
    if (tl == 0) {
      ldsr[k] = rr;
      ldsi[k] = pk;
    }
    // trailing update a_qc -= l_q a_kc over candidate rows x pattern columns
    // (rows past k outside the candidates and columns outside the pattern
    // hold exact zeros: the dense LU's operations on them are skipped)
    for (int pi = tl; pi < (nc - 1) * 8; pi += L) {
      const int cj = 1 + (pi & 7);
      if (cj < np) {
        const int q = cand(k, nbc, 1 + (pi >> 3));
        const int c = patcol(k, nwc, cj);
        const double l = *wp(q, k) * rr;
        const double wk = *wp(k, c);
        double* a = wp(q, c);
        *a = fma(-l, wk, *a);
      }
    }
    // multipliers, stored after every read of column k above
    for (int qi = tl; qi < nc - 1; qi += L) {
      double* a = wp(cand(k, nbc, 1 + qi), k);
      *a = *a * rr;
    }
    team_sync();
  }

  // The substitution chains of band_solve: the factors one position needs,
  // loaded a position ahead, and the chain's running values.  Without an
  // interchange the kT dense rows are updated lane-parallel: lane tl keeps
  // dense row tl & 3 (bdl, its multiplier dl), so their four updates per
  // position are one fma; with interchanges (a position's pivot can be a
  // dense row) every lane keeps all four (bd0..bd3, d0..d3).
  struct FwdFactors {
    int pj;
    double l1, l2, d0, d1, d2, d3, dl, t3;
  };
  struct FwdChain {
    double b0, b1, b2, bd0, bd1, bd2, bd3, bdl;
  };
  struct BackFactors {
    double u6, u5, u4, u3, ri, y;
  };
  struct BackChain {
    double x1, x2, x3, x4;
  };

  // position j's pivot row, multipliers of rows j+1, j+2 and the dense rows,
  // and the right-hand side entry j + 3
  template <bool PIV>
  __device__ __forceinline__ void fwd_load(FwdFactors& f, int j) const {
    if constexpr (PIV) f.pj = ldsi[j];
    f.l1 = Wb[(j + 1) * 8 + 1];
    f.l2 = Wb[(j + 2) * 8];
    if constexpr (PIV) {
      f.d0 = Wd[j];
      f.d1 = Wd[NK + j];
      f.d2 = Wd[2 * NK + j];
      f.d3 = Wd[3 * NK + j];
    } else {
      f.dl = Wd[(tl & 3) * NK + j];
    }
    f.t3 = ldsd[j + 3];
  }

  // PIV: the factorization interchanged rows (else every pj == j).  LAST:
  // position j = nb - 3 = n - 7 (no right-hand side entry j + 3 < nb); every
  // earlier position has one.
  template <bool PIV, bool LAST>
  __device__ __forceinline__ void fwd_step(int j, const FwdFactors& f, FwdFactors& next,
                                           FwdChain& c) {
    const int pj = PIV ? team_uni<L>(f.pj) : j;
    const double bn = LAST ? 0.0 : f.t3;
    fwd_load<PIV>(next, j + 1);
    if constexpr (!PIV) {
      const double xj = c.b0;
      if (tl == 0) ldsd[j] = xj;
      c.b0 = fma(-f.l1, xj, c.b1);
      c.b1 = fma(-f.l2, xj, c.b2);
      c.b2 = bn;
      c.bdl = fma(-f.dl, xj, c.bdl);
      return;
    }
    double b0 = c.b0, b1 = c.b1, b2 = c.b2, bd0 = c.bd0, bd1 = c.bd1, bd2 = c.bd2, bd3 = c.bd3;
    if (PIV && pj != j) {
      double t = b0;
      if (pj == j + 1) {
        b0 = b1;
        b1 = t;
      } else if (pj == j + 2) {
        b0 = b2;
        b2 = t;
      } else if (pj == nb) {
        b0 = bd0;
        bd0 = t;
      } else if (pj == nb + 1) {
        b0 = bd1;
        bd1 = t;
      } else if (pj == nb + 2) {
        b0 = bd2;
        bd2 = t;
      } else {
        b0 = bd3;
        bd3 = t;
      }
    }
    const double xj = b0;
    if (tl == 0) ldsd[j] = xj;
    c.b0 = fma(-f.l1, xj, b1);
    c.b1 = fma(-f.l2, xj, b2);
    c.b2 = bn;
    c.bd0 = fma(-f.d0, xj, bd0);
    c.bd1 = fma(-f.d1, xj, bd1);
    c.bd2 = fma(-f.d2, xj, bd2);
    c.bd3 = fma(-f.d3, xj, bd3);
  }

  // the forward chain over positions 0 .. nb - 3
  template <bool PIV>
  __device__ __forceinline__ void fwd_chain(FwdChain& st) {
    FwdFactors fa, fb;
    fwd_load<PIV>(fa, 0);
    const int jl = nb - 3;
    int j = 0;
    for (; j + 1 < jl; j += 2) {
      fwd_step<PIV, false>(j, fa, fb, st);
      fwd_step<PIV, false>(j + 1, fb, fa, st);
    }
    if (j < jl) {
      fwd_step<PIV, false>(j, fa, fb, st);
      fwd_step<PIV, true>(jl, fb, fa, st);
    } else {
      fwd_step<PIV, true>(jl, fa, fb, st);
    }
  }

  // row i's window entries, reciprocal pivot and right-hand side
  __device__ __forceinline__ void back_load(BackFactors& f, int i) const {
    f.u6 = Wb[i * 8 + 6];
    f.u5 = Wb[i * 8 + 5];
    f.u4 = Wb[i * 8 + 4];
    f.u3 = Wb[i * 8 + 3];
    f.ri = ldsr[i];
    f.y = ldsd[i];
  }

  // row i <= nb - 5: all four window columns are band columns; row 0 loads
  // nothing further
  __device__ __forceinline__ void back_step(int i, const BackFactors& f, BackFactors& next,
                                            BackChain& c) {
    back_load(next, i > 0 ? i - 1 : 0);
    double y = f.y;
    y = fma(-f.u6, c.x4, y);
    y = fma(-f.u5, c.x3, y);
    y = fma(-f.u4, c.x2, y);
    y = fma(-f.u3, c.x1, y);
    const double xi = y * f.ri;
    if (tl == 0) ldsd[i] = xi;
    c.x4 = c.x3;
    c.x3 = c.x2;
    c.x2 = c.x1;
    c.x1 = xi;
  }

  // Solve W x = b with the factors of factor_band; x: owned components in
  // and out.  Forward elimination (interchanges as they come) and back
  // substitution walk the band as chains in registers, every lane of the
  // team computing the same values (lane 0 stores them).  Returns false when
  // a non-finite value met a skipped zero (the dense solve's NaNs): a
  // forward value x_j with j <= nb - 4 (the dense LU multiplies it into the
  // exact zeros of rows past j + 2 that precede the dense ones), or a final
  // x_i with 5 <= i < nb (rows above i - 4 hold exact zeros in column i);
  // both are checked lane-parallel once the chain has stored them.
  __device__ __forceinline__ bool band_solve(double* x) {
    n = team_uni<L>(n);
    nb = team_uni<L>(nb);
#pragma unroll
    for (int m = 0; m < C; m++) {
      int c = comp(m);
      if (c < n) ldsd[c] = x[m];
    }
    team_sync();
    bool bad = false;
    double bd0 = ldsd[nb], bd1 = ldsd[nb + 1], bd2 = ldsd[nb + 2], bd3 = ldsd[nb + 3];
    double b0 = ldsd[0], b1 = ldsd[1], b2 = ldsd[2];
    // forward, positions j with j + 2 < nb: candidates j+1, j+2 and the
    // dense rows (7); rows past j+2 outside them while j < n - 7.  Two
    // register sets of factors alternate: position j + 1's load while
    // position j computes (indices stay inside the team's arrays: j + 4 <=
    // nb + 1 < NK).
    FwdChain st;
    st.b0 = b0;
    st.b1 = b1;
    st.b2 = b2;
    st.bd0 = bd0;
    st.bd1 = bd1;
    st.bd2 = bd2;
    st.bd3 = bd3;
    st.bdl = ldsd[nb + (tl & 3)];
    // without an interchange (the usual case: W = J - I/(gamma dt) is
    // diagonally dominant on this problem) the pivot rows are not read and
    // the dense rows are updated lane-parallel
    if (lu_swapped) {
      fwd_chain<true>(st);
      bd0 = st.bd0;
      bd1 = st.bd1;
      bd2 = st.bd2;
      bd3 = st.bd3;
    } else {
      fwd_chain<false>(st);
      bd0 = __shfl(st.bdl, 0, L);
      bd1 = __shfl(st.bdl, 1, L);
      bd2 = __shfl(st.bdl, 2, L);
      bd3 = __shfl(st.bdl, 3, L);
    }
    b0 = st.b0;
    b1 = st.b1;
#pragma unroll
    for (int m = 0; m < C; m++) {
      const int c = comp(m);
      if (c <= nb - 4) bad = bad || !(fabs(ldsd[c]) <= __DBL_MAX__);
    }
    // the last two band positions and the dense ones: every later position
    // is a candidate
    double v[6] = {b0, b1, bd0, bd1, bd2, bd3};
#pragma unroll
    for (int jj = 0; jj < 6; jj++) {
      const int j = nb - 2 + jj;
      const int pj = ldsi[j] - (nb - 2);
#pragma unroll
      for (int ii = jj + 1; ii < 6; ii++) {
        if (pj == ii) {
          const double t = v[jj];
          v[jj] = v[ii];
          v[ii] = t;
        }
      }
      const double xj = v[jj];
      if (tl == 0) ldsd[j] = xj;
#pragma unroll
      for (int ii = jj + 1; ii < 6; ii++) v[ii] = fma(-*wp(nb - 2 + ii, j), xj, v[ii]);
    }
    // back substitution, dense block (every row above holds its entry)
#pragma unroll
    for (int jj = 5; jj >= 2; jj--) {
      const int j = nb - 2 + jj;
      v[jj] = v[jj] * ldsr[j];
#pragma unroll
      for (int ii = 2; ii < jj; ii++) v[ii] = fma(-Wd[(ii - 2) * NK + j], v[jj], v[ii]);
    }
    if (tl == 0) {
#pragma unroll
      for (int t = 0; t < kT; t++) ldsd[nb + t] = v[2 + t];
    }
    team_sync();
    // band rows: the border columns (j = n-1 down to nb), lanes in parallel
#pragma unroll
    for (int m = 0; m < C; m++) {
      const int r = rowc(m);
      if (r < nb) {
        double y = ldsd[r];
        const double* R = Wr + r * kT;
        y = fma(-R[3], v[5], y);
        y = fma(-R[2], v[4], y);
        y = fma(-R[1], v[3], y);
        y = fma(-R[0], v[2], y);
        ldsd[r] = y;
      }
    }
    team_sync();
    // band rows: the window columns i+4 .. i+1 (those < nb), a chain; the
    // last four rows (nb >= 12) hold fewer window columns
    double x1 = 0.0, x2 = 0.0, x3 = 0.0, x4 = 0.0;
#pragma unroll
    for (int ii = 1; ii <= 4; ii++) {
      const int i = nb - ii;
      const double* U = Wb + i * 8;
      double y = ldsd[i];
      if (ii >= 4) y = fma(-U[5], x3, y);
      if (ii >= 3) y = fma(-U[4], x2, y);
      if (ii >= 2) y = fma(-U[3], x1, y);
      const double xi = y * ldsr[i];
      if (tl == 0) ldsd[i] = xi;
      x4 = x3;
      x3 = x2;
      x2 = x1;
      x1 = xi;
    }
    // the other rows, two register sets alternating (row i - 1's factors
    // load while row i computes)
    {
      BackChain bs;
      bs.x1 = x1;
      bs.x2 = x2;
      bs.x3 = x3;
      bs.x4 = x4;
      BackFactors ba, bb;
      int i = nb - 5;  // >= 7
      back_load(ba, i);
      for (; i - 1 >= 5; i -= 2) {
        back_step(i, ba, bb, bs);
        back_step(i - 1, bb, ba, bs);
      }
      if (i >= 5) {  // i == 5
        back_step(5, ba, bb, bs);
        back_step(4, bb, ba, bs);
        back_step(3, ba, bb, bs);
        back_step(2, bb, ba, bs);
        back_step(1, ba, bb, bs);
        back_step(0, bb, ba, bs);
      } else {  // i == 4
        back_step(4, ba, bb, bs);
        back_step(3, bb, ba, bs);
        back_step(2, ba, bb, bs);
        back_step(1, bb, ba, bs);
        back_step(0, ba, bb, bs);
      }
    }
    team_sync();
#pragma unroll
    for (int m = 0; m < C; m++) {
      const int c = comp(m);
      x[m] = (c < n) ? ldsd[c] : 0.0;
      if (c >= 5 && c < nb) bad = bad || !(fabs(x[m]) <= __DBL_MAX__);
    }
    team_sync();
    return !team_or<L>(bad ? 1 : 0);
  }

  // SAVE builds: Kout receives the dense-output combinations K_m = sum_j
  // H_mj k_j (m < 3) of the step.
  __device__ __forceinline__ bool rodas5_step(double dts, double* u1, double& EEst, double abstol,
                              double reltol, double (*Kout)[C]) {
    const int fac = factor_band(dts);
    if (fac == 1) {
      EEst = 2.0;
      return false;
    }
    if (fac == 2) {
      EEst = __builtin_nan("");
      return true;
    }
    bool solved = true;
    double inv_dt = 1.0 / dts;
    double kd[8][C];
    double U[C], du[C], x[C];
    float Uf[C], duf[C];
#pragma unroll
    for (int s = 0; s < 8; s++) {
      if (!MIX) {
        if (s == 0) {
#pragma unroll
          for (int m = 0; m < C; m++) U[m] = u[m];
        } else if (s <= 5) {
#pragma unroll
          for (int m = 0; m < C; m++) {
            double acc = u[m];
#pragma unroll
            for (int j = 0; j < s; j++) acc = fma(rd_A<kP>(s, j), kd[j][m], acc);
            U[m] = acc;
          }
        } else {
#pragma unroll
          for (int m = 0; m < C; m++) U[m] = U[m] + kd[s - 1][m];
        }
        rhs64(du, U);
      } else {
        if (s == 0) {
#pragma unroll
          for (int m = 0; m < C; m++) Uf[m] = (float)u[m];
        } else if (s <= 5) {
#pragma unroll
          for (int m = 0; m < C; m++) {
            float acc = (float)u[m];
#pragma unroll
            for (int j = 0; j < s; j++) acc = fmaf((float)rd_A<kP>(s, j), (float)kd[j][m], acc);
            Uf[m] = acc;
          }
        } else {
#pragma unroll
          for (int m = 0; m < C; m++) Uf[m] = Uf[m] + (float)kd[s - 1][m];
        }
        rhs32(duf, Uf);
#pragma unroll
        for (int m = 0; m < C; m++) du[m] = (double)duf[m];
      }
#pragma unroll
      for (int m = 0; m < C; m++) {
        double r;
        if (s == 0) {
          r = du[m];
        } else if (!MIX) {
          double acc = (rd_C<kP>(s, 0) * inv_dt) * kd[0][m];
#pragma unroll
          for (int j = 1; j < s; j++) acc = fma(rd_C<kP>(s, j) * inv_dt, kd[j][m], acc);
          r = du[m] + acc;
        } else {
          float idt = (float)inv_dt;
          float acc = ((float)rd_C<kP>(s, 0) * idt) * (float)kd[0][m];
#pragma unroll
          for (int j = 1; j < s; j++) acc = fmaf((float)rd_C<kP>(s, j) * idt, (float)kd[j][m], acc);
          r = (double)((float)du[m] + acc);
        }
        x[m] = -r;
      }
      solved = band_solve(x) && solved;
#pragma unroll
      for (int m = 0; m < C; m++) kd[s][m] = MIX ? (double)(float)x[m] : x[m];
    }
#pragma unroll
    for (int m = 0; m < C; m++) {
      double acc = u[m];
#pragma unroll
      for (int j = 0; j < 5; j++) acc = fma(rd_A<kP>(5, j), kd[j][m], acc);
      acc = acc + kd[5][m];
      acc = acc + kd[6][m];
      u1[m] = acc + kd[7][m];
    }
    EEst = solved ? err_norm(kd[7], u1, abstol, reltol) : __builtin_nan("");
    if constexpr (SAVE) {
#pragma unroll
      for (int m = 0; m < C; m++) {
#pragma unroll
        for (int mm = 0; mm < 3; mm++) {
          double acc = rd_H<kP>(mm, 0) * kd[0][m];
#pragma unroll
          for (int j = 1; j < 7; j++) acc = fma(rd_H<kP>(mm, j), kd[j][m], acc);
          Kout[mm][m] = acc;
        }
      }
    }
    return true;
  }

  // ---------------------------------------------------------------- step
  __device__ __forceinline__ void step(const BrussArgs& a) {
    if (kAuto && iter > 0) {
      double stiffness = fabs(eig * dt / kStabSize);
      bool stiff = stiffness > kStiffTol;
      count = stiff ? (count < 0 ? 1 : count + 1) : (count > 0 ? -1 : count - 1);
      if (mode == 0 && count > kMaxStiff) {
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
    if (iter > maxit) {
      rc = RC_MAXITERS;
      return;
    }
    dt = fmin(dt, dtmax);
    if (a.dtmin > 0 && dt < a.dtmin) {
      rc = RC_DTLESSTHANMIN;
      return;
    }
    dt = fmin(dt, tdir * (tf - t));
    const double sdt = tdir * dt;  // the signed step

    double u1[C];
    double EEst = 0.0;
    double k1save[C];
    float k7f[C];
    // SAVE builds: the step's stages for the dense output
    double ks[SAVE && !MIX ? 7 : 1][C];
    float ksf[SAVE && MIX ? 7 : 1][C];
    double Kr[SAVE ? 3 : 1][C];
    bool tsit_step = kHasTsit5 && (!kHasRodas || mode == 0);
    if (tsit_step) {
      if constexpr (kHasTsit5) {
#pragma unroll
        for (int m = 0; m < C; m++) k1save[m] = k1[m];
        if constexpr (!MIX)
          EEst = tsit5_step64(sdt, u1, a.abstol, a.reltol, ks);
        else
          EEst = tsit5_step_mix(sdt, u1, k7f, a.abstol, a.reltol, ksf);
      }
      nf += 6;
    } else {
      if constexpr (kHasRodas) {
        if (!rodas5_step(sdt, u1, EEst, a.abstol, a.reltol, Kr)) EEst = 2.0;
      }
      nf += 8;
    }
    if (EEst != EEst) {
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
      double ttmp;
      if (!MIX) {
        ttmp = t + sdt;
      } else {
        double y = sdt - tcomp;
        ttmp = t + y;
        tcomp = (ttmp - t) - y;
      }
      const double snapw = tdir > 0.0 ? tsnap : 100.0 * jl_eps(fmax(t, tf));
      if (fabs(ttmp - tf) < snapw) ttmp = tf;
      const double tprev = t;
      t = ttmp;
      if constexpr (SAVE) {
        // savevalues!: every saveat point in (tprev, t] (backward [t, tprev)),
        // from the step's dense output (Tsit5's interpolant or the Rosenbrock
        // one); tnext = +-inf once every point is saved
        while (tdir * (t - tnext) >= 0.0) {
          double val[C];
          if (tnext == t) {
#pragma unroll
            for (int m = 0; m < C; m++) val[m] = u1[m];
          } else {
            const double th = (tnext - tprev) / sdt;
            if (tsit_step) {
              double bt[7];
              tsit5_bthetas(th, bt);
              if constexpr (!MIX) {
#pragma unroll
                for (int m = 0; m < C; m++) {
                  double acc = bt[6] * ks[6][m];
#pragma unroll
                  for (int j = 5; j >= 0; j--) acc = fma(bt[j], ks[j][m], acc);
                  val[m] = fma(sdt, acc, u[m]);
                }
              } else {
                float bf[7];
#pragma unroll
                for (int j = 0; j < 7; j++) bf[j] = (float)bt[j];
#pragma unroll
                for (int m = 0; m < C; m++) {
                  float acc = bf[6] * ksf[6][m];
#pragma unroll
                  for (int j = 5; j >= 0; j--) acc = fmaf(bf[j], ksf[j][m], acc);
                  val[m] = fma(sdt, (double)acc, u[m]);
                }
              }
            } else {
              const double th1 = 1.0 - th;
#pragma unroll
              for (int m = 0; m < C; m++) {
                double inner = fma(th, Kr[2][m], Kr[1][m]);
                inner = fma(th, inner, Kr[0][m]);
                const double tt = fma(th1, inner, u1[m]);
                val[m] = fma(th, tt, th1 * u[m]);
              }
            }
          }
          store_save(a, isave, val);
          isave += tdir > 0.0 ? 1 : -1;
          set_tnext(a);
        }
      }
#pragma unroll
      for (int m = 0; m < C; m++) u[m] = u1[m];
      if (MIX && tsit_step) {
#pragma unroll
        for (int m = 0; m < C; m++) k1f[m] = k7f[m];
      }
      int bad = 0;
#pragma unroll
      for (int m = 0; m < C; m++)
        if (comp(m) < n) bad |= (u1[m] != u1[m]) ? 1 : 0;
      if (team_or<L>(bad)) {
        rc = RC_UNSTABLE;
        return;
      }
      dt = fmin(dtmax, dtnew);
    } else {
      nreject++;
      if (tsit_step && !MIX) {
#pragma unroll
        for (int m = 0; m < C; m++) k1[m] = k1save[m];
      }
      if constexpr (!FAST)
        dt = dt / fmin(1.0 / kQmin, q11 / kGammaSafety);
      else
        dt = dt * fast_exp2(-fmin(kL2QminInv, l11 - kL2Gamma));
    }
    if (!(tdir * (tf - t) > 0.0)) rc = RC_SUCCESS;
  }

  __device__ __forceinline__ void finish(const BrussArgs& a) {
    const int64_t NT = a.ntraj;
    if constexpr (SAVE) {
      // points past the end of an unfinished solve
      double nanv[C];
#pragma unroll
      for (int m = 0; m < C; m++) nanv[m] = __builtin_nan("");
      if (tdir > 0.0)
        for (int k = isave; k < a.ns; k++) store_save(a, k, nanv);
      else
        for (int k = isave; k >= 0; k--) store_save(a, k, nanv);
    }
    if (a.u_final)
#pragma unroll
      for (int m = 0; m < C; m++) {
        int c = comp(m);
        if (c < n) a.u_final[(int64_t)c * NT + i] = u[m];
      }
    if (tl == 0) {
      if (a.t_final) a.t_final[i] = t;
      if (a.retcode) a.retcode[i] = rc;
      if (a.naccept) a.naccept[i] = naccept;
      if (a.nreject) a.nreject[i] = nreject;
      if (a.nf) a.nf[i] = nf;
      if (a.nstiff) a.nstiff[i] = nstiff;
      if (a.dt_first) a.dt_first[i] = dtfirst;
    }
  }
};

// NK: largest state size the instantiation holds (C = NK / L slots per
// lane); every instance it is given must have n <= NK.
template <int L, int NK, bool MIX, int ALG, bool FAST, bool SAVE>
__device__ __forceinline__ void bruss_run(const BrussArgs& a) {
  constexpr int C = NK / L;
  constexpr int TPB = 256 / L;
  using Team = BTeam<L, C, MIX, ALG, FAST, SAVE>;
  // One block of LDS per team, every offset inside it small enough for a DS
  // instruction's immediate: the exchange vector and the reciprocal pivots
  // side by side (position i's pair loads as one ds_read2), then W in band +
  // border form (8 NK band entries, kT NK border columns, kT dense rows of
  // NK; Rosenbrock-capable kernels), the FP32 exchange vector (mixed
  // precision) and the pivot rows (Rosenbrock-capable kernels).
  constexpr int kR = Team::kHasRodas ? NK : 0;
  constexpr int kW = Team::kHasRodas ? (8 + 2 * Team::kT) * NK : 0;
  constexpr int kF = MIX ? NK / 2 : 0;
  constexpr int kI = Team::kHasRodas ? NK / 2 : 0;
  constexpr int PER = NK + kR + kW + kF + kI;
  __shared__ double s_team[TPB][PER];
  const int team = threadIdx.x / L;
  Team T;
  T.tl = threadIdx.x % L;
  double* const base = s_team[team];
  T.ldsd = base;
  T.ldsr = base + NK;
  T.ldsf = (float*)(base + NK + kR + kW);
  T.ldsi = (int*)(base + NK + kR + kW + kF);
  if constexpr (Team::kHasRodas) {
    T.Wb = base + NK + kR;
    T.Wr = T.Wb + 8 * NK;
    T.Wd = T.Wr + Team::kT * NK;
  }
  bool active = false, drained = false;
  while (true) {
    if (!active && !drained) {
      unsigned long long q = 0;
      if (T.tl == 0) q = atomicAdd(a.queue, 1ull);
      q = team_uni_u64<L>(__shfl(q, 0, L));
      if (q >= (unsigned long long)a.nqueue) {
        drained = true;
      } else {
        int64_t idx = (int64_t)team_uni_u64<L>(
            (unsigned long long)(a.order ? a.order[q] : (int64_t)q));
        // an order entry outside [0, ntraj) is skipped
        if ((uint64_t)idx < (uint64_t)a.ntraj) {
          T.init(a, idx);
          if (T.rc == RC_DEFAULT)
            active = true;
          else
            T.finish(a);
        }
      }
    }
    if (!__any(active || !drained)) break;
    if (active) {
      T.step(a);
      if (T.rc != RC_DEFAULT) {
        T.finish(a);
        active = false;
      }
    }
  }
}

template <int L, int NK, bool MIX, int ALG, bool FAST, bool SAVE>
__global__ __launch_bounds__(256) void k_bruss(BrussArgs a) {
  bruss_run<L, NK, MIX, ALG, FAST, SAVE>(a);
}

// The 32-lane FP64 Tsit5 team holding two (u, v) pairs per lane: left alone
// the compiler takes 135 VGPRs (3 waves per SIMD, VALU busy ~84 %); held to
// 4 waves it spills 5 dwords and runs 5 % faster (131,072-instance batch:
// 34.3 -> 32.7 ms).  The FP32-stage build fits 4 waves by itself.
template <int L, int NK, bool MIX, bool FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bruss_w4(
    BrussArgs a) {
  bruss_run<L, NK, MIX, ALG_TSIT5, FAST, false>(a);
}

template <int L, int NK, bool MIX, int ALG, bool FAST, bool SAVE>
static constexpr auto bruss_kernel() {
  if constexpr (ALG == ALG_TSIT5 && L == 32 && NK == BRUSS_NMAX && !MIX && !SAVE)
    return k_bruss_w4<L, NK, MIX, FAST>;
  else
    return k_bruss<L, NK, MIX, ALG, FAST, SAVE>;
}

// One instantiation per (team shape, precision, algorithm, controller),
// with or without saveat (SAVE: instantiated in kernels_bruss_save.hip).
template <bool SAVE>
struct OccupancyOp {
  int* per_cu;
  template <int L, int NK, bool MIX, int ALG, bool FAST>
  hipError_t run() const {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        per_cu, bruss_kernel<L, NK, MIX, ALG, FAST, SAVE>(), 256, 0);
  }
};

template <bool SAVE>
struct LaunchOp {
  const BrussArgs* a;
  int grid;
  hipStream_t s;
  template <int L, int NK, bool MIX, int ALG, bool FAST>
  hipError_t run() const {
    hipLaunchKernelGGL((bruss_kernel<L, NK, MIX, ALG, FAST, SAVE>()), dim3(grid), dim3(256), 0,
                       s, *a);
    return hipGetLastError();
  }
};

template <int L, int NK, bool MIX, bool FAST, class Op>
static hipError_t by_alg(int alg, const Op& op) {
  switch (alg) {
    case ALG_TSIT5: return op.template run<L, NK, MIX, ALG_TSIT5, FAST>();
    case ALG_RODAS5: return op.template run<L, NK, MIX, ALG_RODAS5, FAST>();
    case ALG_RODAS5P: return op.template run<L, NK, MIX, ALG_RODAS5P, FAST>();
    case ALG_AUTO5P: return op.template run<L, NK, MIX, ALG_AUTO5P, FAST>();
    default: return op.template run<L, NK, MIX, ALG_AUTO, FAST>();
  }
}

template <int L, int NK, class Op>
static hipError_t by_prec(int alg, int mix, int fast, const Op& op) {
  if (mix) return fast ? by_alg<L, NK, true, true>(alg, op) : by_alg<L, NK, true, false>(alg, op);
  return fast ? by_alg<L, NK, false, true>(alg, op) : by_alg<L, NK, false, false>(alg, op);
}

// Team shapes: 64 or 32 lanes holding up to 128 components, or 32 lanes
// holding up to 64 (two instances per wavefront, one (u, v) pair per lane:
// the registers of the 64-lane shape for batches with n <= 64).
template <class Op>
static hipError_t dispatch(int lanes, int nk, int alg, int mix, int fast, const Op& op) {
  if (lanes == 64 && nk == BRUSS_NMAX) return by_prec<64, BRUSS_NMAX>(alg, mix, fast, op);
  if (lanes == 32 && nk == BRUSS_NMAX) return by_prec<32, BRUSS_NMAX>(alg, mix, fast, op);
  if (lanes == 32 && nk == BRUSS_NMAX / 2) return by_prec<32, BRUSS_NMAX / 2>(alg, mix, fast, op);
  return hipErrorInvalidValue;
}

}  // namespace mideq
