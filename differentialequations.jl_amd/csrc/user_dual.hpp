// user_dual.hpp -- forward-mode dual numbers for the Jacobian of run-time
// compiled right-hand sides (OrdinaryDiffEq's default for a Rosenbrock
// method without a user Jacobian is forward-mode automatic differentiation).
//
// A user RHS is a template over its number type T; instantiated with
// Dual<N>, whose value parts follow the double arithmetic exactly, each
// partial carries d/du_j, so J[i][j] = du[i].d[j].  Every derivative rule is
// written out with one fixed operation order.  This text is compiled as
// device code (hipRTC) and, by the tests, as host code; MIDEQ_HD sets the
// function qualifiers.
#ifndef MIDEQ_HD
#define MIDEQ_HD __device__ __forceinline__
#endif

namespace mideq {

// the double and float overloads next to the Dual ones below
using ::cos;
using ::exp;
using ::fabs;
using ::log;
using ::pow;
using ::sin;
using ::sqrt;
using ::tanh;

template <int K>
struct Dual {
  double v;
  double d[K];
  MIDEQ_HD Dual() : v(0.0) {
    for (int j = 0; j < K; j++) d[j] = 0.0;
  }
  MIDEQ_HD Dual(double x) : v(x) {
    for (int j = 0; j < K; j++) d[j] = 0.0;
  }
};

// (a + b)' = a' + b'
template <int K>
MIDEQ_HD Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r(a.v + b.v);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] + b.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator+(const Dual<K>& a, double c) {
  Dual<K> r(a.v + c);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator+(double c, const Dual<K>& a) {
  Dual<K> r(c + a.v);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r(-a.v);
  for (int j = 0; j < K; j++) r.d[j] = -a.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r(a.v - b.v);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] - b.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator-(const Dual<K>& a, double c) {
  Dual<K> r(a.v - c);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator-(double c, const Dual<K>& a) {
  Dual<K> r(c - a.v);
  for (int j = 0; j < K; j++) r.d[j] = -a.d[j];
  return r;
}
// (a b)' = a' b + a b'
template <int K>
MIDEQ_HD Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r(a.v * b.v);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator*(const Dual<K>& a, double c) {
  Dual<K> r(a.v * c);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] * c;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator*(double c, const Dual<K>& a) {
  Dual<K> r(c * a.v);
  for (int j = 0; j < K; j++) r.d[j] = c * a.d[j];
  return r;
}
// (a / b)' = (a' b - a b') / (b b)
template <int K>
MIDEQ_HD Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r(a.v / b.v);
  double bb = b.v * b.v;
  for (int j = 0; j < K; j++) r.d[j] = (a.d[j] * b.v - a.v * b.d[j]) / bb;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator/(const Dual<K>& a, double c) {
  Dual<K> r(a.v / c);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] / c;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> operator/(double c, const Dual<K>& b) {
  Dual<K> r(c / b.v);
  double bb = b.v * b.v;
  for (int j = 0; j < K; j++) r.d[j] = -(c * b.d[j]) / bb;
  return r;
}
template <int K>
MIDEQ_HD Dual<K>& operator+=(Dual<K>& a, const Dual<K>& b) { return a = a + b; }
template <int K>
MIDEQ_HD Dual<K>& operator-=(Dual<K>& a, const Dual<K>& b) { return a = a - b; }
template <int K>
MIDEQ_HD Dual<K>& operator*=(Dual<K>& a, const Dual<K>& b) { return a = a * b; }
template <int K>
MIDEQ_HD Dual<K>& operator/=(Dual<K>& a, const Dual<K>& b) { return a = a / b; }
template <int K>
MIDEQ_HD Dual<K>& operator+=(Dual<K>& a, double c) { return a = a + c; }
template <int K>
MIDEQ_HD Dual<K>& operator-=(Dual<K>& a, double c) { return a = a - c; }
template <int K>
MIDEQ_HD Dual<K>& operator*=(Dual<K>& a, double c) { return a = a * c; }
template <int K>
MIDEQ_HD Dual<K>& operator/=(Dual<K>& a, double c) { return a = a / c; }

// comparisons look at values (branches in a RHS)
template <int K>
MIDEQ_HD bool operator<(const Dual<K>& a, const Dual<K>& b) { return a.v < b.v; }
template <int K>
MIDEQ_HD bool operator>(const Dual<K>& a, const Dual<K>& b) { return a.v > b.v; }
template <int K>
MIDEQ_HD bool operator<=(const Dual<K>& a, const Dual<K>& b) { return a.v <= b.v; }
template <int K>
MIDEQ_HD bool operator>=(const Dual<K>& a, const Dual<K>& b) { return a.v >= b.v; }
template <int K>
MIDEQ_HD bool operator<(const Dual<K>& a, double c) { return a.v < c; }
template <int K>
MIDEQ_HD bool operator>(const Dual<K>& a, double c) { return a.v > c; }
template <int K>
MIDEQ_HD bool operator<(double c, const Dual<K>& a) { return c < a.v; }
template <int K>
MIDEQ_HD bool operator>(double c, const Dual<K>& a) { return c > a.v; }

// elementary functions: f(a)' = f'(a.v) a'
template <int K>
MIDEQ_HD Dual<K> sqrt(const Dual<K>& a) {
  double s = ::sqrt(a.v);
  Dual<K> r(s);
  double den = 2.0 * s;
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] / den;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> exp(const Dual<K>& a) {
  double e = ::exp(a.v);
  Dual<K> r(e);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] * e;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> log(const Dual<K>& a) {
  Dual<K> r(::log(a.v));
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] / a.v;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> sin(const Dual<K>& a) {
  Dual<K> r(::sin(a.v));
  double c = ::cos(a.v);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] * c;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> cos(const Dual<K>& a) {
  Dual<K> r(::cos(a.v));
  double s = ::sin(a.v);
  for (int j = 0; j < K; j++) r.d[j] = -(a.d[j] * s);
  return r;
}
template <int K>
MIDEQ_HD Dual<K> tanh(const Dual<K>& a) {
  double th = ::tanh(a.v);
  Dual<K> r(th);
  double g = 1.0 - th * th;
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] * g;
  return r;
}
template <int K>
MIDEQ_HD Dual<K> fabs(const Dual<K>& a) {
  return a.v < 0.0 ? -a : a;
}
template <int K>
MIDEQ_HD Dual<K> pow(const Dual<K>& a, double c) {
  Dual<K> r(::pow(a.v, c));
  double g = c * ::pow(a.v, c - 1.0);
  for (int j = 0; j < K; j++) r.d[j] = a.d[j] * g;
  return r;
}

}  // namespace mideq
