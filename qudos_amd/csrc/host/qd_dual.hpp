// Forward-mode dual numbers for exact first derivatives of the host coefficient code with respect
// to one shape parameter: Dual carries (value, d value / d theta) of a real quantity, CDual the
// same for a complex one.  Branch tests compare values only, so a dual evaluation follows the same
// branch as the plain double evaluation and returns its value unchanged.
#pragma once

#include <cmath>
#include <complex>

namespace qd {

struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double value) : v(value) {}   // NOLINT: constants promote implicitly
    Dual(double value, double deriv) : v(value), d(deriv) {}
};

inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator/(double a, Dual b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
inline Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
inline Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
inline Dual &operator*=(Dual &a, Dual b) { a = a * b; return a; }

inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline double der(double) { return 0.0; }
inline double der(Dual x) { return x.d; }

inline Dual sqrt(Dual a) {
    const double s = std::sqrt(a.v);
    return {s, s > 0.0 ? 0.5 * a.d / s : 0.0};
}
inline Dual sin(Dual a) { return {std::sin(a.v), std::cos(a.v) * a.d}; }
inline Dual cos(Dual a) { return {std::cos(a.v), -std::sin(a.v) * a.d}; }
inline Dual exp(Dual a) { const double e = std::exp(a.v); return {e, e * a.d}; }
inline Dual atan(Dual a) { return {std::atan(a.v), a.d / (1.0 + a.v * a.v)}; }
inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
// Bessel functions of the first kind: J0' = -J1, J1' = J0 - J1 / x
inline Dual j0(Dual a) { return {::j0(a.v), -::j1(a.v) * a.d}; }
inline Dual j1(Dual a) {
    const double j1v = ::j1(a.v);
    const double dj1 = a.v != 0.0 ? ::j0(a.v) - j1v / a.v : 0.5;
    return {j1v, dj1 * a.d};
}

// Complex dual: value and derivative are complex.
struct CDual {
    std::complex<double> v, d;
    CDual() = default;
    CDual(std::complex<double> value) : v(value) {}   // NOLINT
    CDual(double value) : v(value, 0.0) {}            // NOLINT
    CDual(Dual re) : v(re.v, 0.0), d(re.d, 0.0) {}    // NOLINT
    CDual(std::complex<double> value, std::complex<double> deriv) : v(value), d(deriv) {}
    CDual(Dual re, Dual im) : v(re.v, im.v), d(re.d, im.d) {}
};

inline CDual operator+(const CDual &a, const CDual &b) { return {a.v + b.v, a.d + b.d}; }
inline CDual operator-(const CDual &a, const CDual &b) { return {a.v - b.v, a.d - b.d}; }
inline CDual operator-(const CDual &a) { return {-a.v, -a.d}; }
inline CDual operator*(const CDual &a, const CDual &b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline CDual operator/(const CDual &a, const CDual &b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline CDual &operator+=(CDual &a, const CDual &b) { a = a + b; return a; }
inline CDual &operator-=(CDual &a, const CDual &b) { a = a - b; return a; }
inline CDual &operator*=(CDual &a, const CDual &b) { a = a * b; return a; }
// mixed with real duals and plain complex numbers
inline CDual operator*(const CDual &a, Dual b) { return a * CDual(b); }
inline CDual operator*(Dual a, const CDual &b) { return CDual(a) * b; }
inline CDual operator/(const CDual &a, Dual b) { return a / CDual(b); }
inline CDual operator/(Dual a, const CDual &b) { return CDual(a) / b; }
inline CDual operator+(const CDual &a, Dual b) { return a + CDual(b); }
inline CDual operator+(Dual a, const CDual &b) { return CDual(a) + b; }
inline CDual operator-(const CDual &a, Dual b) { return a - CDual(b); }
inline CDual operator-(Dual a, const CDual &b) { return CDual(a) - b; }
inline CDual operator*(std::complex<double> a, Dual b) { return {a * b.v, a * b.d}; }
inline CDual operator*(Dual a, std::complex<double> b) { return {a.v * b, a.d * b}; }
inline CDual operator/(std::complex<double> a, Dual b) { return CDual(a) / CDual(b); }
inline CDual operator+(std::complex<double> a, Dual b) { return {a + b.v, std::complex<double>(b.d, 0.0)}; }
inline CDual operator-(std::complex<double> a, Dual b) { return {a - b.v, std::complex<double>(-b.d, 0.0)}; }
inline CDual operator-(Dual a, std::complex<double> b) { return {a.v - b, std::complex<double>(a.d, 0.0)}; }
inline CDual operator+(Dual a, std::complex<double> b) { return {a.v + b, std::complex<double>(a.d, 0.0)}; }
// with plain doubles and complex numbers (explicit, so that no call is ambiguous between the
// Dual and the CDual conversions of a double)
inline CDual operator+(const CDual &a, double b) { return {a.v + b, a.d}; }
inline CDual operator+(double a, const CDual &b) { return {a + b.v, b.d}; }
inline CDual operator-(const CDual &a, double b) { return {a.v - b, a.d}; }
inline CDual operator-(double a, const CDual &b) { return {a - b.v, -b.d}; }
inline CDual operator*(const CDual &a, double b) { return {a.v * b, a.d * b}; }
inline CDual operator*(double a, const CDual &b) { return {a * b.v, a * b.d}; }
inline CDual operator/(const CDual &a, double b) { return {a.v / b, a.d / b}; }
inline CDual operator/(double a, const CDual &b) { return CDual(std::complex<double>(a, 0.0)) / b; }
inline CDual operator+(const CDual &a, std::complex<double> b) { return {a.v + b, a.d}; }
inline CDual operator+(std::complex<double> a, const CDual &b) { return {a + b.v, b.d}; }
inline CDual operator-(const CDual &a, std::complex<double> b) { return {a.v - b, a.d}; }
inline CDual operator-(std::complex<double> a, const CDual &b) { return {a - b.v, -b.d}; }
inline CDual operator*(const CDual &a, std::complex<double> b) { return {a.v * b, a.d * b}; }
inline CDual operator*(std::complex<double> a, const CDual &b) { return {a * b.v, a * b.d}; }
inline CDual operator/(const CDual &a, std::complex<double> b) { return {a.v / b, a.d / b}; }
inline CDual operator/(std::complex<double> a, const CDual &b) { return CDual(a) / b; }
inline CDual exp(const CDual &a) { const std::complex<double> e = std::exp(a.v); return {e, e * a.d}; }
inline double abs(const CDual &a) { return std::abs(a.v); }

inline std::complex<double> val(std::complex<double> x) { return x; }
inline std::complex<double> val(const CDual &x) { return x.v; }
inline std::complex<double> der(std::complex<double>) { return 0.0; }
inline std::complex<double> der(const CDual &x) { return x.d; }

// complex scalar type that goes with a real scalar type
template <class R> struct CplxOf;
template <> struct CplxOf<double> { using type = std::complex<double>; };
template <> struct CplxOf<Dual> { using type = CDual; };

}  // namespace qd
