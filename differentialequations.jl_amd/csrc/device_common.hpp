// device_common.hpp -- method constants, controller arithmetic and models
// shared by the MI355X ensemble kernels.
//
// Everything here is evaluated per lane in registers.  The FP64 arithmetic
// follows one fixed convention (explicit fma where OrdinaryDiffEq's @muladd
// fuses, everything else separately rounded; the library is compiled with
// -ffp-contract=off), so an FP64 solve is bit-reproducible and can be
// compared exactly, step count for step count, with a CPU restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mideq {

// ---------------------------------------------------------------------------
// Tsit5 (Tsitouras 2011; OrdinaryDiffEq Tsit5ConstantCacheActual).
// ---------------------------------------------------------------------------
namespace ts {
constexpr double c2 = 0.161, c3 = 0.327, c4 = 0.9, c5 = 0.9800255409045097;
constexpr double a21 = 0.161;
constexpr double a31 = -0.008480655492356989, a32 = 0.335480655492357;
constexpr double a41 = 2.897153057105493, a42 = -6.359448489975075,
                 a43 = 4.3622954328695815;
constexpr double a51 = 5.325864828439257, a52 = -11.748883564062828,
                 a53 = 7.4955393428898365, a54 = -0.09249506636175525;
constexpr double a61 = 5.86145544294642, a62 = -12.92096931784711,
                 a63 = 8.159367898576159, a64 = -0.071584973281401,
                 a65 = -0.028269050394068383;
constexpr double a71 = 0.09646076681806523, a72 = 0.01,
                 a73 = 0.4798896504144996, a74 = 1.379008574103742,
                 a75 = -3.290069515436081, a76 = 2.324710524099774;
constexpr double bt1 = -0.00178001105222577714, bt2 = -0.0008164344596567469,
                 bt3 = 0.007880878010261995, bt4 = -0.1447110071732629,
                 bt5 = 0.5823571654525552, bt6 = -0.45808210592918697,
                 bt7 = 0.015151515151515152;
// dense output polynomials b_i(theta)
constexpr double r11 = 1.0, r12 = -2.763706197274826, r13 = 2.9132554618219126,
                 r14 = -1.0530884977290216;
constexpr double r22 = 0.13169999999999998, r23 = -0.2234, r24 = 0.1017;
constexpr double r32 = 3.9302962368947516, r33 = -5.941033872131505,
                 r34 = 2.490627285651253;
constexpr double r42 = -12.411077166933676, r43 = 30.33818863028232,
                 r44 = -16.548102889244902;
constexpr double r52 = 37.50931341651104, r53 = -88.1789048947664,
                 r54 = 47.37952196281928;
constexpr double r62 = -27.896526289197286, r63 = 65.09189467479366,
                 r64 = -34.87065786149661;
constexpr double r72 = 1.5, r73 = -4.0, r74 = 2.5;
}  // namespace ts

// ---------------------------------------------------------------------------
// Rodas5 (Di Marzo 1993 / Hairer-Wanner II), transformed W-form.
// ---------------------------------------------------------------------------
namespace rd {
constexpr double gamma = 0.19;
constexpr double A[6][5] = {
    {0, 0, 0, 0, 0},
    {2.0, 0, 0, 0, 0},
    {3.040894194418781, 1.041747909077569, 0, 0, 0},
    {2.576417536461461, 1.622083060776640, -0.9089668560264532, 0, 0},
    {2.760842080225597, 1.446624659844071, -0.3036980084553738,
     0.2877498600325443, 0},
    {-14.09640773051259, 6.925207756232704, -41.47510893210728,
     2.343771018586405, 24.13215229196062}};
constexpr double C[8][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {-10.31323885133993, 0, 0, 0, 0, 0, 0},
    {-21.04823117650003, -7.234992135176716, 0, 0, 0, 0, 0},
    {32.22751541853323, -4.943732386540191, 19.44922031041879, 0, 0, 0, 0},
    {-20.69865579590063, -8.816374604402768, 1.260436877740897,
     -0.7495647613787146, 0, 0, 0},
    {-46.22004352711257, -17.49534862857472, -289.6389582892057,
     93.60855400400906, 318.3822534212147, 0, 0},
    {34.20013733472935, -14.15535402717690, 57.82335640988400,
     25.83362985412365, 1.408950972071624, -6.551835421242162, 0},
    {42.57076742291101, -13.80770672017997, 93.98938432427124,
     18.77919633714503, -31.58359187223370, -6.685968952921985,
     -5.810979938412932}};
// stage times c_i (row sums of the A coefficients of the stage inputs) and
// the time-derivative weights d_i (row sums of Gamma), used by
// non-autonomous right-hand sides
constexpr double c[8] = {0.0, 0.38, 0.3878509998321533, 0.4839718937873840,
                         0.4570477008819580, 1.0, 1.0, 1.0};
constexpr double d[8] = {0.19, -0.1823079225333714636, -0.319231832186874912,
                         0.3449828624712811, -0.377417564392089818, 0.0, 0.0, 0.0};
// Continuous extension (this library's derivation; see rd_H below).
constexpr double H[3][7] = {
    {46.335061333323615, -6.136914790748748, 108.40712618588684, -15.432459674306823,
     -79.5099979954734, -1.9050165705652515, -2.7727981784318065},
    {-39.14106358101113, -3.4608988577808986, -157.10746284880244, 34.66137042565768,
     146.53644774500066, 4.746961424327032, 8.03303264503124},
    {35.36683737689737, 3.300351318514294, 162.0764418691076, -9.977463628946106,
     -129.9992347337098, -8.753639789000294, -10.993114152187328}};
}  // namespace rd

// ---------------------------------------------------------------------------
// Rodas5P (Steinebach, BIT 63 (2023) 27): same W-form, 8 stages and stiffly
// accurate structure as Rodas5, new coefficients.
// ---------------------------------------------------------------------------
namespace rd5p {
constexpr double gamma = 0.21193756319429014;
constexpr double A[6][5] = {
    {0, 0, 0, 0, 0},
    {3.0, 0, 0, 0, 0},
    {2.849394379747939, 0.45842242204463923, 0, 0, 0},
    {-6.954028509809101, 2.489845061869568, -10.358996098473584, 0, 0},
    {2.8029986275628964, 0.5072464736228206, -0.3988312541770524, -0.04721187230404641, 0},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946}};
constexpr double C[8][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {-14.155112264123755, 0, 0, 0, 0, 0, 0},
    {-17.97296035885952, -2.859693295451294, 0, 0, 0, 0, 0},
    {147.12150275711716, -1.41221402718213, 71.68940251302358, 0, 0, 0, 0},
    {165.43517024871676, -0.4592823456491126, 42.90938336958603, -5.961986721573306, 0, 0, 0},
    {24.854864614690072, -3.0009227002832186, 47.4931110020768, 5.5814197821558125,
     -0.6610691825249471, 0, 0},
    {30.91273214028599, -3.1208243349937974, 77.79954646070892, 34.28646028294783,
     -19.097331116725623, -28.087943162872662, 0},
    {37.80277123390563, -3.2571969029072276, 112.26918849496327, 66.9347231244047,
     -40.06618937091002, -54.66780262877968, -9.48861652309627}};
constexpr double c[8] = {0.0, 0.6358126895828704, 0.4095798393397535, 0.9769306725060716,
                         0.4288403609558664, 1.0, 1.0, 1.0};
constexpr double d[8] = {0.21193756319429014, -0.42387512638858027, -0.3384627126235924,
                         1.8046452872882734, 2.325825639765069, 0.0, 0.0, 0.0};
constexpr double H[3][7] = {
    {18.16834780539935, -2.4039764856508183, -28.49033949791586, -39.2353239248467,
     24.203612973337464, 31.138884328789313, 5.813992390809251},
    {20.204639179440537, -1.56515746765803, 153.73045429293606, 118.22912892648397,
     -71.49081351328653, -100.21520495678435, -27.642530934779916},
    {-15.086356891234155, 3.4125953268012186, -59.68473410632594, -45.47407592791441,
     28.595178703267447, 42.46386356049424, 18.195502534440195}};
}  // namespace rd5p

// Rosenbrock tableau selected at compile time (P: Rodas5P, else Rodas5).
template <bool P>
__device__ __forceinline__ constexpr double rd_gamma() { return P ? rd5p::gamma : rd::gamma; }
template <bool P>
__device__ __forceinline__ constexpr double rd_A(int s, int j) {
  return P ? rd5p::A[s][j] : rd::A[s][j];
}
template <bool P>
__device__ __forceinline__ constexpr double rd_C(int s, int j) {
  return P ? rd5p::C[s][j] : rd::C[s][j];
}
template <bool P>
__device__ __forceinline__ constexpr double rd_c(int s) { return P ? rd5p::c[s] : rd::c[s]; }
template <bool P>
__device__ __forceinline__ constexpr double rd_d(int s) { return P ? rd5p::d[s] : rd::d[s]; }
// Dense output of a Rosenbrock step (saveat inside a Rodas5/Rodas5P step):
//   u(th) = (1-th) uprev + th (u1 + (1-th)(K1 + th (K2 + th K3))),
//   K_m = sum_{j<7} H[m][j] k_j
// with k_j the stage increments.  H is the unique solution, for each m, of
// (a) the eight Rosenbrock order conditions up to order 4 holding for every
// th (so the interpolant is locally 4th order), (b) the stiff limit
// u(th)/uprev -> (1-th)^4 for y' = lambda y as lambda dt -> -inf, and (c) no
// weight on the error stage k_8.  (a) has rank 6 for these tableaus;
// (a)+(b)+(c) is a consistent 9 x 7 system of rank 7.
template <bool P>
__device__ __forceinline__ constexpr double rd_H(int m, int j) {
  return P ? rd5p::H[m][j] : rd::H[m][j];
}

// ---------------------------------------------------------------------------
// Controller constants (PI controller defaults for an order-5 method) and
// the auto-switching thresholds of AutoTsit5(Rodas5()).
// ---------------------------------------------------------------------------
constexpr double kBeta1 = 7.0 / 50.0, kBeta2 = 2.0 / 25.0, kGammaSafety = 0.9;
constexpr double kQmin = 0.2, kQmax = 10.0, kQoldInit = 1e-4;
constexpr double kStabSize = 3.5068;
constexpr int kMaxStiff = 10, kMaxNonStiff = 3;
constexpr double kStiffTol = 0.9, kDtFac = 2.0;
constexpr double kStiffTolS = kStiffTol * kStabSize;  // fast form of the stiffness test

enum : int8_t { RC_DEFAULT = 0, RC_SUCCESS = 1, RC_MAXITERS = 2, RC_DTLESSTHANMIN = 3,
                RC_UNSTABLE = 4, RC_DTNAN = 5, RC_INITIALFAILURE = 6, RC_TERMINATED = 7,
                // internal: the first pass of a two-phase AutoSwitch solve hands
                // the trajectory to the full AutoSwitch pass (never returned)
                RC_HANDOFF = 100 };

// FP64 constant pinned to an SGPR pair at its point of use.  gfx9 VOP3
// instructions take one scalar operand and no literal; an unpinned constant
// addend of a Horner step is first copied into a VGPR pair (two v_mov_b32 on
// the vector ALU) so that the fma can be shrunk to v_fmac.  Pinned, it costs
// two scalar moves and the fma stays one VOP3 instruction.  Same value, same
// operation: results are unchanged.
__device__ __forceinline__ double kpin(double c) {
  asm volatile("" : "+s"(c));
  return c;
}

// ---------------------------------------------------------------------------
// Controller power x^y = exp2(y log2 x): atanh series for ln of the mantissa
// in [sqrt(1/2), sqrt(2)), Taylor series of 2^f for |f| <= 1/2.  Relative
// error ~1e-11, far below what step control needs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ctl_log2(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = fma(s2, 1.0 / 13.0, 1.0 / 11.0);
  p = fma(s2, p, kpin(1.0 / 9.0));
  p = fma(s2, p, kpin(1.0 / 7.0));
  p = fma(s2, p, kpin(1.0 / 5.0));
  p = fma(s2, p, kpin(1.0 / 3.0));
  p = fma(s2, p, 1.0);
  double lnm = (2.0 * s) * p;
  return fma(lnm, 1.4426950408889634074, (double)e);
}

__device__ __forceinline__ double ctl_exp2(double z) {
  if (z != z) return z;
  if (z > 1020.0) return __builtin_inf();
  if (z < -1020.0) return 0.0;
  double nn = rint(z);
  double f = z - nn;
  double x = f * 0.69314718055994530942;
  double p = fma(x, 1.0 / 362880.0, 1.0 / 40320.0);
  p = fma(x, p, kpin(1.0 / 5040.0));
  p = fma(x, p, kpin(1.0 / 720.0));
  p = fma(x, p, kpin(1.0 / 120.0));
  p = fma(x, p, kpin(1.0 / 24.0));
  p = fma(x, p, kpin(1.0 / 6.0));
  p = fma(x, p, 0.5);
  p = fma(x, p, 1.0);
  p = fma(x, p, 1.0);
  return ldexp(p, (int)nn);
}

__device__ __forceinline__ double ctl_pow(double x, double y) {
  return ctl_exp2(y * ctl_log2(x));
}

// Division-free log2 for the fast controller: log2(1 + t) = t * q(t) on
// t = m - 1, m in [sqrt(1/2), sqrt(2)), q a degree-8 least-squares fit at
// Chebyshev nodes (|error| < 1.4e-8).
__device__ __forceinline__ double fast_log2(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double t = m - 1.0;
  double q = 0.12614846624505;
  q = fma(q, t, -0.20742103505566906);
  q = fma(q, t, kpin(0.2156698589589577));
  q = fma(q, t, kpin(-0.23892033981765542));
  q = fma(q, t, kpin(0.28791832447651855));
  q = fma(q, t, kpin(-0.36070482809673765));
  q = fma(q, t, kpin(0.4809106099449815));
  q = fma(q, t, kpin(-0.7213473333598942));
  q = fma(q, t, kpin(1.442695003764732));
  return fma(t, q, (double)e);
}

// 2^z for the fast controller's step factors, whose exponent is clamped to
// |z| <= log2(10): z = n + f, n = rint(z), |f| <= 1/2, degree-6 Taylor
// polynomial of e^x at x = f ln 2 (|x| <= 0.347, relative error < 2e-7,
// exact at integer z).  No range checks: the domain is bounded.
__device__ __forceinline__ double fast_exp2(double z) {
  double nn = rint(z);
  double f = z - nn;
  double x = f * 0.69314718055994530942;
  double p = fma(x, 1.0 / 720.0, 1.0 / 120.0);
  p = fma(x, p, kpin(1.0 / 24.0));
  p = fma(x, p, kpin(1.0 / 6.0));
  p = fma(x, p, 0.5);
  p = fma(x, p, 1.0);
  p = fma(x, p, 1.0);
  return ldexp(p, (int)nn);
}

// log2 of the controller constants (fast form)
constexpr double kL2QmaxInv = -3.321928094887362;   // log2(1/10)
constexpr double kL2QminInv = 2.321928094887362;    // log2(5)
constexpr double kL2Gamma = -0.15200309344504997;   // log2(0.9)
constexpr double kL2QoldInit = -13.287712379549449; // log2(1e-4)
constexpr double kL2QsteadyRodas = 0.2630344058337938;  // log2(6/5)

// Julia eps(x): spacing to the next larger float.
__device__ __forceinline__ double jl_eps(double x) {
  x = fabs(x);
  return nextafter(x, __builtin_inf()) - x;
}

// ---------------------------------------------------------------------------
// Models.  Each provides f in FP64 and FP32 (same operation order), an
// analytic Jacobian in FP64 (row-major J[i][j]) where it has one, and the
// number of states/parameters.  The operation order follows each RHS as
// written in Julia (left-to-right binary operators, x^2 as x*x).
// ---------------------------------------------------------------------------
struct ModelExp {
  static constexpr int N = 1, NP = 1;
  template <class T>
  __device__ static void f(T* du, const T* u, const T* p) { du[0] = p[0] * u[0]; }
  __device__ static void jac(double (*J)[N], const double* u, const double* p) {
    (void)u;
    J[0][0] = p[0];
  }
};

struct ModelLorenz {
  static constexpr int N = 3, NP = 3;
  template <class T>
  __device__ static void f(T* du, const T* u, const T* p) {
    du[0] = p[0] * (u[1] - u[0]);
    du[1] = u[0] * (p[1] - u[2]) - u[1];
    du[2] = u[0] * u[1] - p[2] * u[2];
  }
  __device__ static void jac(double (*J)[N], const double* u, const double* p) {
    J[0][0] = -p[0]; J[0][1] = p[0]; J[0][2] = 0.0;
    J[1][0] = p[1] - u[2]; J[1][1] = -1.0; J[1][2] = -u[0];
    J[2][0] = u[1]; J[2][1] = u[0]; J[2][2] = -p[2];
  }
};

struct ModelRobertson {
  static constexpr int N = 3, NP = 3;
  template <class T>
  __device__ static void f(T* du, const T* u, const T* p) {
    T y1 = u[0], y2 = u[1], y3 = u[2];
    T k1 = p[0], k2 = p[1], k3 = p[2];
    du[0] = -k1 * y1 + k3 * y2 * y3;
    du[1] = k1 * y1 - k2 * (y2 * y2) - k3 * y2 * y3;
    du[2] = k2 * (y2 * y2);
  }
  __device__ static void jac(double (*J)[N], const double* u, const double* p) {
    double y2 = u[1], y3 = u[2], k1 = p[0], k2 = p[1], k3 = p[2];
    J[0][0] = -k1; J[0][1] = k3 * y3; J[0][2] = k3 * y2;
    J[1][0] = k1; J[1][1] = -2.0 * k2 * y2 - k3 * y3; J[1][2] = -k3 * y2;
    J[2][0] = 0.0; J[2][1] = 2.0 * k2 * y2; J[2][2] = 0.0;
  }
};

struct ModelVdP {
  static constexpr int N = 2, NP = 1;
  template <class T>
  __device__ static void f(T* du, const T* u, const T* p) {
    T x = u[0], y = u[1], mu = p[0];
    du[0] = y;
    du[1] = mu * (T(1) - x * x) * y - x;
  }
  __device__ static void jac(double (*J)[N], const double* u, const double* p) {
    double x = u[0], y = u[1], mu = p[0];
    J[0][0] = 0.0; J[0][1] = 1.0;
    J[1][0] = -2.0 * mu * x * y - 1.0; J[1][1] = mu * (1.0 - x * x);
  }
};

struct ModelLinear3 {
  static constexpr int N = 3, NP = 9;
  template <class T>
  __device__ static void f(T* du, const T* u, const T* p) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      T acc = p[i * 3] * u[0];
      acc = fma_t(p[i * 3 + 1], u[1], acc);
      acc = fma_t(p[i * 3 + 2], u[2], acc);
      du[i] = acc;
    }
  }
  __device__ static void jac(double (*J)[N], const double* u, const double* p) {
    (void)u;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) J[i][j] = p[i * 3 + j];
  }
  __device__ static double fma_t(double a, double b, double c) { return fma(a, b, c); }
  __device__ static float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
};

}  // namespace mideq
