// moist_dual.h — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// First-order dual numbers for the oracle's templates: a value and one
// tangent, propagated by the chain rule through every arithmetic operation
// and elementary function the oracle uses.  Comparisons, min/max, signbit and
// isnan look at the value only, so every branch follows the primal
// trajectory: instantiating the oracle's rk3_step with FwdAD is forward-mode
// differentiation of the restated reference with its branches frozen, what
// ForwardDiff.jl does through moist.jl (dual numbers with the same rules).
// The value part is computed by exactly the double operations the double
// instantiation performs, so the primal trajectory is the double oracle's.
#pragma once

#include <cmath>

namespace orc {

struct FwdAD {
    double v = 0.0, d = 0.0;
    FwdAD() = default;
    FwdAD(double x) : v(x), d(0.0) {}  // constants and parameters: zero tangent
    FwdAD(double x, double t) : v(x), d(t) {}
    explicit operator double() const { return v; }
    FwdAD &operator+=(const FwdAD &b) { v += b.v; d += b.d; return *this; }
    FwdAD &operator-=(const FwdAD &b) { v -= b.v; d -= b.d; return *this; }
    FwdAD &operator*=(const FwdAD &b) { *this = FwdAD(v * b.v, d * b.v + v * b.d); return *this; }
    FwdAD &operator/=(const FwdAD &b) { *this = FwdAD(v / b.v, (d * b.v - v * b.d) / (b.v * b.v)); return *this; }
};
inline FwdAD operator+(const FwdAD &a, const FwdAD &b) { return FwdAD(a.v + b.v, a.d + b.d); }
inline FwdAD operator-(const FwdAD &a, const FwdAD &b) { return FwdAD(a.v - b.v, a.d - b.d); }
inline FwdAD operator-(const FwdAD &a) { return FwdAD(-a.v, -a.d); }
inline FwdAD operator*(const FwdAD &a, const FwdAD &b) { return FwdAD(a.v * b.v, a.d * b.v + a.v * b.d); }
inline FwdAD operator/(const FwdAD &a, const FwdAD &b) {
    const double q = a.v / b.v;
    return FwdAD(q, (a.d - q * b.d) / b.v);
}
inline bool operator<(const FwdAD &a, const FwdAD &b) { return a.v < b.v; }
inline bool operator>(const FwdAD &a, const FwdAD &b) { return a.v > b.v; }
inline bool operator<=(const FwdAD &a, const FwdAD &b) { return a.v <= b.v; }
inline bool operator>=(const FwdAD &a, const FwdAD &b) { return a.v >= b.v; }
inline bool operator==(const FwdAD &a, const FwdAD &b) { return a.v == b.v; }
inline bool operator!=(const FwdAD &a, const FwdAD &b) { return a.v != b.v; }

// The Float64 off-centring of the acoustic explicit part (moist.jl:1375):
// a + fact_tp * inc evaluated in Float64 and converted back to T.
template <class T> inline T off_centre(T a, double f, T inc) { return T(double(a) + f * double(inc)); }
inline FwdAD off_centre(FwdAD a, double f, FwdAD inc) { return FwdAD(a.v + f * inc.v, a.d + f * inc.d); }

}  // namespace orc

// Elementary functions of a FwdAD, declared in std so that the oracle's
// qualified calls (std::exp, ...) find them.
namespace std {
inline orc::FwdAD exp(const orc::FwdAD &a) {
    const double e = std::exp(a.v);
    return orc::FwdAD(e, e * a.d);
}
inline orc::FwdAD sqrt(const orc::FwdAD &a) {
    const double r = std::sqrt(a.v);
    return orc::FwdAD(r, a.d / (2.0 * r));
}
inline orc::FwdAD tanh(const orc::FwdAD &a) {
    const double t = std::tanh(a.v);
    return orc::FwdAD(t, (1.0 - t * t) * a.d);
}
inline orc::FwdAD fabs(const orc::FwdAD &a) { return std::signbit(a.v) ? -a : a; }
// x^y: d = y x^(y-1) dx + x^y log(x) dy (the log term only when y varies)
inline orc::FwdAD pow(const orc::FwdAD &x, const orc::FwdAD &y) {
    const double p = std::pow(x.v, y.v);
    double t = y.v * std::pow(x.v, y.v - 1.0) * x.d;
    if (y.d != 0.0) t += p * std::log(x.v) * y.d;
    return orc::FwdAD(p, t);
}
inline bool isnan(const orc::FwdAD &a) { return std::isnan(a.v); }
inline bool signbit(const orc::FwdAD &a) { return std::signbit(a.v); }
}  // namespace std
