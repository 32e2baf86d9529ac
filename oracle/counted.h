// CPU ORACLE -- test infrastructure only (never linked by the product).
//
// An operation-counting scalar: the oracle's FP64 code instantiated with
// Counted instead of double computes the same values and counts the
// floating-point operations it performs, which gives the algorithmic work per
// node and RHS evaluation: +, -, * count 1 each; /, sqrt, log, exp, pow count
// 1 each (their own counter); comparisons, fabs, fmin/fmax and negation count 0.
// Counting runs are single-threaded (the counters are plain globals).
#ifndef ORA_COUNTED_H
#define ORA_COUNTED_H

#include <cmath>

struct OraCounts {
  long long add = 0, mul = 0, div = 0, special = 0;
};
OraCounts& ora_counts();

struct Counted {
  double v = 0.0;
  Counted() = default;
  Counted(double x) : v(x) {}  // NOLINT: implicit, so literals mix in
  explicit operator double() const { return v; }
  Counted& operator+=(Counted o) { ++ora_counts().add; v += o.v; return *this; }
  Counted& operator-=(Counted o) { ++ora_counts().add; v -= o.v; return *this; }
  Counted& operator*=(Counted o) { ++ora_counts().mul; v *= o.v; return *this; }
  Counted& operator/=(Counted o) { ++ora_counts().div; v /= o.v; return *this; }
};

inline Counted operator+(Counted a, Counted b) { ++ora_counts().add; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { ++ora_counts().add; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { ++ora_counts().mul; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { ++ora_counts().div; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
inline Counted sqrt(Counted a) { ++ora_counts().special; return Counted(std::sqrt(a.v)); }
inline Counted log(Counted a) { ++ora_counts().special; return Counted(std::log(a.v)); }
inline Counted exp(Counted a) { ++ora_counts().special; return Counted(std::exp(a.v)); }
inline Counted pow(Counted a, Counted b) { ++ora_counts().special; return Counted(std::pow(a.v, b.v)); }
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted fmax(Counted a, Counted b) { return Counted(std::fmax(a.v, b.v)); }
inline Counted fmin(Counted a, Counted b) { return Counted(std::fmin(a.v, b.v)); }

// the same names for double, so templated oracle code calls them unqualified
using std::exp;
using std::fabs;
using std::fmax;
using std::fmin;
using std::log;
using std::pow;
using std::sqrt;

// Logarithmic mean with the f^2 < 1e-4 series branch, Ranocha et al. (2023),
// "Efficient implementation of modern entropy stable and kinetic energy
// preserving discontinuous Galerkin methods for conservation laws", ACM TOMS,
// Section 4 (the form Trixi's ln_mean implements; imported src/TrixiAtmo.jl:37)
template <typename R>
R ora_ln_mean_t(R x, R y) {
  const R f2 = (x * (x - 2 * y) + y * y) / (x * (x + 2 * y) + y * y);
  if (f2 < 1.0e-4) {
    return (x + y) / (2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0))));
  }
  return (y - x) / log(y / x);
}

template <typename R>
R ora_inv_ln_mean_t(R x, R y) {
  const R f2 = (x * (x - 2 * y) + y * y) / (x * (x + 2 * y) + y * y);
  if (f2 < 1.0e-4) {
    return (2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0)))) / (x + y);
  }
  return log(y / x) / (y - x);
}

#endif
