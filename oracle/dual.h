/* Forward-mode dual numbers for the oracle's tangent build (test
 * infrastructure only: built with -DORACLE_DUAL into liboracle_dual.so).
 *
 * A Dual carries a value v and a directional derivative d.  Arithmetic is
 * the value's FP64 arithmetic (so the value of every quantity is the one the
 * plain oracle computes, bit for bit) plus the chain rule for d.  Decisions
 * (comparisons, fmin/fmax, floor to a cell index) use the value alone, and
 * the derivative follows the branch taken.  The FP32 length keeps the
 * reference's FP32 value and takes the derivative of the exact root, the
 * linearisation the engine's adjoint uses.
 */
#ifndef ORACLE_DUAL_H_
#define ORACLE_DUAL_H_

#include <math.h>

struct Dual {
    double v, d;
    Dual() : v(0.0), d(0.0) {}
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants and inputs lift implicitly
    Dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    Dual& operator+=(const Dual& b) { d = d + b.d; v = v + b.v; return *this; }
    Dual& operator-=(const Dual& b) { d = d - b.d; v = v - b.v; return *this; }
    Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v = v * b.v; return *this; }
    Dual& operator/=(const Dual& b) {
        const double q = v / b.v;
        d = (d - q * b.d) / b.v;
        v = q;
        return *this;
    }
};

static inline Dual operator+(Dual a, const Dual& b) { return a += b; }
static inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
static inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
static inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
static inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
static inline Dual operator+(const Dual& a) { return a; }

static inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
static inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
static inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
static inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
static inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
static inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

static inline Dual sqrt(const Dual& a) {
    const double r = ::sqrt(a.v);
    return Dual(r, r > 0.0 ? a.d / (2.0 * r) : 0.0);
}
static inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
static inline Dual fmin(const Dual& a, const Dual& b) { return b.v < a.v ? b : a; }
static inline Dual fmax(const Dual& a, const Dual& b) { return b.v > a.v ? b : a; }
static inline Dual sin(const Dual& a) { return Dual(::sin(a.v), ::cos(a.v) * a.d); }
static inline Dual cos(const Dual& a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }
static inline Dual exp(const Dual& a) {
    const double e = ::exp(a.v);
    return Dual(e, e * a.d);
}
static inline Dual pow(const Dual& a, double p) {
    const double r = ::pow(a.v, p);
    return Dual(r, a.v != 0.0 ? p * ::pow(a.v, p - 1.0) * a.d : 0.0);
}
static inline double floor(const Dual& a) { return ::floor(a.v); }
static inline bool isnan(const Dual& a) { return a.v != a.v; }

#endif
