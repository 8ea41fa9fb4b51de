// Shape transforms chi_s(G) = (1/Omega) int_s exp(-i G.r) d^3r of all 19 shape types on the device
// (characteristic_function_fourier.c:58-1105 with the pyramid and prism segments of :1112-1396 and
// the 1-D integrals of characteristic_function_fourier_integrals.c), with the Euler rotation, the
// centre's phase and the Gaussian softening.  The shapes of revolution with curved tops (hemisphere,
// lens, cones, truncated cones and rings, 3, 4, 11-15) need 1-D integrals of Bessel functions: they
// use the host's Gauss-Legendre pair (40 and 64 points, interval halving to 1e-14, shapes.cpp),
// visiting the subintervals in the host recursion's order without a stack.
//
// One work-item per wave vector and one work-group row per shape (blockIdx.y): a graded structure
// of many slices (the Truncated_pyramid example: 100 pyramid slices on a 49 x 49 x 81 strain grid;
// the QD molecule: 12 truncated-cone slices) is one launch.  The branches of the closed forms
// (G = 0, G along an axis, G in a plane) are taken per work-item on the rotated wave vector, as on
// the host.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "qd_host.hpp"
#include "qd_internal.hpp"
#include "qd_shapes_device.hpp"

namespace qd {

namespace {

constexpr double kPiS = 3.14159265358979323846;

// ---- real scalars: double, or a forward-mode dual number (value, derivative with respect to one
// shape length) for the exact derivatives of the transforms.  The helpers carry an x prefix so that
// the double overloads of the math library stay visible.
struct Dd {
    double v, d;
};
__device__ inline Dd operator+(Dd a, Dd b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dd operator-(Dd a, Dd b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dd operator-(Dd a) { return {-a.v, -a.d}; }
__device__ inline Dd operator*(Dd a, Dd b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dd operator/(Dd a, Dd b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ inline Dd operator+(Dd a, double b) { return {a.v + b, a.d}; }
__device__ inline Dd operator+(double a, Dd b) { return {a + b.v, b.d}; }
__device__ inline Dd operator-(Dd a, double b) { return {a.v - b, a.d}; }
__device__ inline Dd operator-(double a, Dd b) { return {a - b.v, -b.d}; }
__device__ inline Dd operator*(Dd a, double b) { return {a.v * b, a.d * b}; }
__device__ inline Dd operator*(double a, Dd b) { return {a * b.v, a * b.d}; }
__device__ inline Dd operator/(Dd a, double b) { return {a.v / b, a.d / b}; }
__device__ inline Dd operator/(double a, Dd b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
__device__ inline cpx<Dd> operator*(double s, cpx<Dd> a) { return {s * a.re, s * a.im}; }

__device__ inline double val(double x) { return x; }
__device__ inline double val(Dd x) { return x.v; }
__device__ inline double xsin(double a) { return sin(a); }
__device__ inline Dd xsin(Dd a) { return {sin(a.v), cos(a.v) * a.d}; }
__device__ inline double xcos(double a) { return cos(a); }
__device__ inline Dd xcos(Dd a) { return {cos(a.v), -sin(a.v) * a.d}; }
__device__ inline void xsincos(double a, double &s, double &c) { sincos(a, &s, &c); }
__device__ inline void xsincos(Dd a, Dd &s, Dd &c) {
    double sv, cv;
    sincos(a.v, &sv, &cv);
    s = {sv, cv * a.d};
    c = {cv, -sv * a.d};
}
__device__ inline double xsqrt(double a) { return sqrt(a); }
__device__ inline Dd xsqrt(Dd a) {
    const double s = sqrt(a.v);
    return {s, s > 0.0 ? 0.5 * a.d / s : 0.0};
}
__device__ inline double xatan(double a) { return atan(a); }
__device__ inline Dd xatan(Dd a) { return {atan(a.v), a.d / (1.0 + a.v * a.v)}; }
__device__ inline double xj0(double a) { return j0(a); }
__device__ inline Dd xj0(Dd a) { return {j0(a.v), -j1(a.v) * a.d}; }
__device__ inline double xj1(double a) { return j1(a); }
__device__ inline Dd xj1(Dd a) {
    const double j1v = j1(a.v);
    return {j1v, (a.v != 0.0 ? j0(a.v) - j1v / a.v : 0.5) * a.d};
}
template <class R> __device__ inline R seeded(double v, double d);
template <> __device__ inline double seeded<double>(double v, double) { return v; }
template <> __device__ inline Dd seeded<Dd>(double v, double d) { return {v, d}; }

template <class R> using CR = cpx<R>;
template <class R> __device__ inline CR<R> cis_m(R a) {   // e^{-ia}
    R s, c;
    xsincos(a, s, c);
    return mk(c, -s);
}
template <class R> __device__ inline CR<R> expi(R a) {    // e^{ia}
    R s, c;
    xsincos(a, s, c);
    return mk(c, s);
}
template <class R> __device__ inline CR<R> ci(R x) { return mk(R{0.0}, x); }         // i x
template <class R> __device__ inline CR<R> times_i(CR<R> z) { return mk(-z.im, z.re); }   // i z
template <class R> __device__ inline CR<R> cr(R x) { return mk(x, R{0.0}); }
template <class R> __device__ inline CR<R> divr(CR<R> z, R r) { return (R{1.0} / r) * z; }
template <class R> __device__ inline CR<R> divd(CR<R> z, double r) { return (1.0 / r) * z; }
template <class R> __device__ inline CR<R> addr(CR<R> z, double r) { return mk(z.re + r, z.im); }
template <class R> __device__ inline CR<R> lift(c64 z) { return mk(R{z.re}, R{z.im}); }

template <class R> __device__ inline R sinc_len(double G, R L) {   // L sin(GL/2)/(GL/2)
    const R t = 0.5 * G * L;
    return val(t) == 0.0 ? L : L * xsin(t) / t;
}
template <class R> __device__ inline R sphere_kernel(R g) { return (xsin(g) - g * xcos(g)) / (g * g * g); }

// (e^{iGa} - 1)/(iG), a at G = 0
template <class R> __device__ inline CR<R> edge_integral(R a, R G) {
    if (fabs(val(G)) < kEps) return cr(a);
    return div_(addr(expi(G * a), -1.0), ci(G));
}

// Triangular pyramid segment k: apex height h over the base plane, base edge at x0 with half-width
// y0.  nside == 0: rectangular base of half lengths L1, L2, the segment turned by 2 atan(L2/L1) k
// or 2 atan(L1/L2) k; nside > 0: regular polygon of apothem L1, turned by 2 pi k / nside.
template <class R>
__device__ CR<R> pyramid_segment(double Gx0, double Gy0, double Gz, int k, R L1, R L2, R h, int nside) {
    R alpha, x0, y0;
    if (nside > 0) { alpha = R{kPiS / nside}; x0 = L1; y0 = tan(kPiS / nside) * L1; }
    else if (k == 0 || k == 2) { alpha = xatan(L2 / L1); x0 = L1; y0 = L2; }
    else { alpha = xatan(L1 / L2); x0 = L2; y0 = L1; }
    const R ang = 2.0 * alpha * (double)k;
    const R Gx = xcos(ang) * Gx0 + xsin(ang) * Gy0;
    const R Gy = -xsin(ang) * Gx0 + xcos(ang) * Gy0;
    const R Gm = xsqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const bool zx = fabs(val(Gx)) < kEps, zy = fabs(val(Gy)) < kEps, zz = fabs(Gz) < kEps;
    if (val(Gm) < kEps) return cr(x0 * y0 * h / 3.0);
    if (zx && zy) {
        const CR<R> t = divr(addr(ci(-Gz * h) - cis_m(Gz * h), 1.0), Gz * Gz * h * h);
        return times_i(cr(2.0 * y0 * x0 / Gz)) * addr(t, -0.5);
    }
    if (zx && zz) return cr((2.0 * h * x0 / (Gy * Gy * y0)) * (1.0 - xsin(Gy * y0) / (Gy * y0)));
    if (zy && zz) {
        const CR<R> e = cis_m(Gx * x0);
        const CR<R> v = addr(-e, -1.0) - ci(2.0 / (Gx * x0)) + times_i(divr(mk(2.0 * e.re, 2.0 * e.im), Gx * x0));
        return (2.0 * h * y0 / (Gx * Gx * x0)) * v;
    }
    if (zx) {
        const R sp = Gy * y0 + Gz * h, sm = Gy * y0 - Gz * h;
        const CR<R> pre = ci(-x0 / (Gz * Gy));
        const CR<R> tail = cr(-2.0 * (xcos(Gy * y0) - 1.0) / (Gy * y0));
        CR<R> body;
        if (fabs(val(sm)) < kEps) body = mk(R{0.0}, R{-1.0}) + divr(addr(expi(sp), -1.0), sp);
        else if (fabs(val(sp)) < kEps) body = divr(addr(cis_m(sm), -1.0), sm) + mk(R{0.0}, R{1.0});
        else body = divr(addr(cis_m(sm), -1.0), sm) + divr(addr(expi(sp), -1.0), sp);
        return pre * (cis_m(Gz * h) * body + tail);
    }
    if (zy) {
        const R q = Gx - Gz * h / x0;
        const CR<R> base = divr(addr(cis_m(Gx * x0) * addr(ci(Gx * x0), 1.0), -1.0), Gx * Gx);
        const CR<R> pre = ci(2.0 * y0 / (Gz * x0));
        if (fabs(val(q)) < kEps) return pre * ((0.5 * x0 * x0) * cis_m(Gz * h) - base);
        const R r = Gx * x0 - Gz * h;
        return pre * (cis_m(Gz * h) * divr(addr(cis_m(r) * addr(ci(r), 1.0), -1.0), q * q) - base);
    }
    if (zz) {
        const R pp = Gx + Gy * y0 / x0, pm = Gx - Gy * y0 / x0;
        const CR<R> Ap = ci(1.0 / (x0 * pp * pp)) * addr(cis_m(x0 * pp), -1.0);
        const CR<R> Am = ci(1.0 / (x0 * pm * pm)) * addr(cis_m(x0 * pm), -1.0);
        if (fabs(val(pm)) < kEps) return (h / Gy) * (cr(1.0 / pp) - Ap - ci(0.5 * x0));
        if (fabs(val(pp)) < kEps) return (h / Gy) * (cr(-1.0 / pm) + Am + ci(0.5 * x0));
        return (h / Gy) * (cr(1.0 / pp - 1.0 / pm) - Ap + Am);
    }
    const R s1 = Gy * y0 / x0, s2 = Gz * h / x0;
    const CR<R> ez = cis_m(Gz * h);
    return (R{-1.0} / (Gy * Gz)) * (edge_integral(x0, -Gx - s1 + s2) * ez - edge_integral(x0, -Gx + s1 + s2) * ez -
                                    edge_integral(x0, -Gx - s1) + edge_integral(x0, -Gx + s1));
}

// Wedge segment k of a regular N-sided prism of inradius x0 and height h centred in z.
template <class R>
__device__ CR<R> prism_segment(double Gx0, double Gy0, double Gz, int k, int N, R x0, R h) {
    const double alpha = kPiS / N, ang = 2.0 * kPiS * k / N;
    const double Gx = cos(ang) * Gx0 + sin(ang) * Gy0;
    const double Gy = -sin(ang) * Gx0 + cos(ang) * Gy0;
    const double Gm = sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const double ta = tan(alpha);
    const bool zx = fabs(Gx) < kEps, zy = fabs(Gy) < kEps, zz = fabs(Gz) < kEps;
    if (Gm < kEps) return cr(x0 * x0 * h * ta);
    if (zx && zy) return cr(2.0 * x0 * x0 * ta * xsin(0.5 * Gz * h) / Gz);
    if (zx && zz) return cr((2.0 * h / (Gy * Gy * ta)) * (1.0 - xcos(Gy * x0 * ta)));
    const CR<R> ex = cis_m(Gx * x0);
    const CR<R> xterm = ci(x0) * ex + divd(addr(ex, -1.0), Gx);
    if (zy && zz) return (2.0 * h * ta / Gx) * xterm;
    if (zx) return cr((4.0 * xsin(0.5 * Gz * h) / (Gy * Gy * Gz * ta)) * (1.0 - xcos(Gy * x0 * ta)));
    if (zy) return (4.0 * xsin(0.5 * Gz * h) * ta / (Gz * Gx)) * xterm;
    const double up = Gy * ta + Gx, um = -Gy * ta + Gx;
    const CR<R> ep = addr(-cis_m(Gy * x0 * ta + Gx * x0), 1.0);
    const CR<R> em = addr(-cis_m(-Gy * x0 * ta + Gx * x0), 1.0);
    CR<R> face;
    if (fabs(up) < kEps) face = ci(x0) - divd(em, um);
    else if (fabs(um) < kEps) face = divd(ep, up) - ci(x0);
    else face = divd(ep, up) - divd(em, um);
    if (zz) return (h / Gy) * face;
    return ((2.0 / (Gy * Gz)) * xsin(0.5 * Gz * h)) * face;
}

// ---- 1-D integrals: the host's Gauss-Legendre pair (40 and 64 points) with interval halving until
// the two agree to 1e-14 (depth at most 12), the subintervals visited in the host recursion's order
// without a stack.  gl holds x40, w40, x64, w64.
template <class R, class F>
__device__ CR<R> gl_apply(const F &f, const double *x, const double *w, int n, R a, R b) {
    const R h = 0.5 * (b - a), c = 0.5 * (b + a);
    CR<R> s = cr(R{0.0});
    for (int i = 0; i < n; ++i) s = s + w[i] * f(c + h * x[i]);
    return h * s;
}
template <class R, class F>
__device__ CR<R> integrate(const F &f, R a, R b, const double *gl) {
    const double *x40 = gl, *w40 = gl + kGLLo, *x64 = gl + 2 * kGLLo, *w64 = x64 + kGLHi;
    // depth-first over the dyadic subintervals, left halves first: interval k of depth d is
    // [a + (b - a) k / 2^d, a + (b - a) (k + 1) / 2^d]; no stack (private arrays would live in scratch)
    CR<R> total = cr(R{0.0});
    int d = 0;
    long k = 0;
    for (;;) {
        const R w = (b - a) * ldexp(1.0, -d);
        const R lo_a = a + w * (double)k, hi_b = a + w * (double)(k + 1);
        const CR<R> lo = gl_apply(f, x40, w40, kGLLo, lo_a, hi_b);
        const CR<R> hi = gl_apply(f, x64, w64, kGLHi, lo_a, hi_b);
        const double hr = val(hi.re), hm = val(hi.im), dr = hr - val(lo.re), dm = hm - val(lo.im);
        if (sqrt(dr * dr + dm * dm) <= 1e-14 * fmax(1.0, sqrt(hr * hr + hm * hm)) || d >= 12) {
            total = total + hi;
            while (k & 1) {   // a right half is done: its parent is too
                k >>= 1;
                --d;
            }
            if (d == 0) break;
            ++k;              // the right sibling
        } else {
            ++d;              // the left half
            k <<= 1;
        }
    }
    return total;
}

// int_0^1 x cos(alpha sqrt(1-x^2)) J0(beta x) dx
template <class R> __device__ R int_J0(R alpha, R beta, const double *gl) {
    auto f = [&](R x) {
        const R u = 1.0 - x * x;
        return cr(x * xcos(alpha * xsqrt(val(u) > 0 ? u : R{0.0})) * xj0(beta * x));
    };
    return integrate(f, R{0.0}, R{1.0}, gl).re;
}
// int over [R1/(R2-R1), R2/(R2-R1)] of x J0((R2-R1) alpha x) e^{i beta x}
template <class R> __device__ CR<R> int_TC(R alpha, R R1, R R2, R beta, const double *gl) {
    auto f = [&](R x) { return (x * xj0((R2 - R1) * alpha * x)) * expi(beta * x); };
    return integrate(f, R1 / (R2 - R1), R2 / (R2 - R1), gl);
}
// int over the same limits of x^2 J0((R2-R1) alpha x)
template <class R> __device__ R int_TC_z(R alpha, R R1, R R2, const double *gl) {
    auto f = [&](R x) { return cr(x * x * xj0((R2 - R1) * alpha * x)); };
    return integrate(f, R1 / (R2 - R1), R2 / (R2 - R1), gl).re;
}
// int_0^1 x J0(alpha x) e^{i beta x} dx
template <class R> __device__ CR<R> int_2(R alpha, R beta, const double *gl) {
    auto f = [&](R x) { return (x * xj0(alpha * x)) * expi(beta * x); };
    return integrate(f, R{0.0}, R{1.0}, gl);
}
// int_0^1 x^2 J0(alpha x) dx
template <class R> __device__ R int_2_z(R alpha, const double *gl) {
    auto f = [&](R x) { return cr(x * x * xj0(alpha * x)); };
    return integrate(f, R{0.0}, R{1.0}, gl).re;
}
// int_fr^1 x J0(alpha x) e^{i beta x / (1 - fr)} dx
template <class R> __device__ CR<R> int_2_etc(R alpha, R beta, R fr, const double *gl) {
    auto f = [&](R x) { return (x * xj0(alpha * x)) * expi(beta * x / (1.0 - fr)); };
    return integrate(f, fr, R{1.0}, gl);
}
// int_fr^1 x^2 J0(alpha x) dx
template <class R> __device__ R int_2_etc_z(R alpha, R fr, const double *gl) {
    auto f = [&](R x) { return cr(x * x * xj0(alpha * x)); };
    return integrate(f, fr, R{1.0}, gl).re;
}

// Cone over a disc (gq = |G_par|, Rr = radius, A = Rr^2) or an ellipse (gq = |(Gx L1, Gy L2)|, Rr = 1,
// A = L1 L2), height H, base on the origin.
template <class R> __device__ CR<R> cone(double Gm, R gq, double Gz, R Rr, R A, R H, const double *gl) {
    if (fabs(Gm) < kEps) return cr(kPiS * A * H / 3.0);
    if (fabs(val(gq)) < kEps) {
        const CR<R> t = divr(addr(ci(-Gz * H) - cis_m(Gz * H), 1.0), (Gz * H) * (Gz * H));
        return ci(2.0 * kPiS * A / Gz) * addr(t, -0.5);
    }
    const R a = gq * Rr;
    if (fabs(Gz) < kEps) return cr(2.0 * kPiS * A * H * (xj1(a) / a - int_2_z(a, gl)));
    return ci(2.0 * kPiS * A / Gz) * (cis_m(Gz * H) * int_2(a, Gz * H, gl)) - ci(2.0 * kPiS * A * xj1(a) / (a * Gz));
}

// Truncated cone: base radius R1 on the origin, top radius R2 at height H.
template <class R> __device__ CR<R> trunc_cone(double Gm, double Gp, double Gz, R R1, R R2, R H, const double *gl) {
    const R dR = R1 - R2;
    if (fabs(Gm) < kEps) return cr(kPiS * H * (R1 * R1 + R1 * R2 + R2 * R2) / 3.0);
    if (fabs(Gp) < kEps) {
        const R a = Gz * H;
        const CR<R> e = cis_m(a);
        const CR<R> v = (0.5 * R2 * R2) * addr(e, -1.0) +
                        ((dR * dR) / (a * a)) * (e * addr(ci(Gz * R2 * H / dR), -1.0) + addr(ci(-R1 * Gz * H / dR), 1.0));
        return ci(R{2.0 * kPiS / Gz}) * (v - cr(0.5 * (R1 * R1 - R2 * R2)));
    }
    const R jj = (R1 * xj1(Gp * R1) - R2 * xj1(Gp * R2)) / Gp;
    if (fabs(Gz) < kEps)
        return cr(2.0 * kPiS * H * (R2 * xj1(Gp * R2) / Gp - dR * dR * int_TC_z(R{Gp}, R2, R1, gl) + (R1 / dR) * jj));
    const CR<R> v = (R2 * xj1(Gp * R2) / Gp) * addr(cis_m(Gz * H), -1.0) +
                    (dR * dR) * (cis_m(Gz * H * R1 / dR) * int_TC(R{Gp}, R2, R1, Gz * H, gl));
    return ci(R{2.0 * kPiS / Gz}) * (v - cr(jj));
}

// Truncated elliptical cone: base semi-axes L1, L2, top L1 fr, L2 fr at height H; gq = |(Gx L1, Gy L2)|.
template <class R>
__device__ CR<R> trunc_econe(double Gm, R gq, double Gz, R L1, R L2, R fr, R H, const double *gl) {
    const R A = L1 * L2;
    if (fabs(Gm) < kEps) return cr(kPiS * A * H * (1.0 + fr + fr * fr) / 3.0);
    const CR<R> e = cis_m(Gz * H);
    if (fabs(val(gq)) < kEps) {
        const CR<R> inner = e * (ci(-Gz * H * fr) + cr(1.0 - fr)) + (ci(Gz * H) + cr(fr - 1.0));
        const CR<R> v = (0.5 * fr * fr) * addr(e, -1.0) + ((fr - 1.0) / (Gz * Gz * H * H)) * inner;
        return ci(2.0 * kPiS * A / Gz) * (v - cr(0.5 * (1.0 - fr * fr)));
    }
    const R jb = (xj1(gq) - fr * xj1(fr * gq)) / gq;
    if (fabs(Gz) < kEps)
        return cr(2.0 * kPiS * A * H * (fr * xj1(fr * gq) / gq - int_2_etc_z(gq, fr, gl) / (1.0 - fr) + jb / (1.0 - fr)));
    const CR<R> v = (fr * xj1(fr * gq) / gq) * addr(e, -1.0) + cis_m(Gz * H / (1.0 - fr)) * int_2_etc(gq, Gz * H, fr, gl);
    return ci(2.0 * kPiS * A / Gz) * (v - cr(jb));
}

// Cylinder of radius Rr and height h centred on its origin.
template <class R> __device__ R cylinder(double Gm, double Gp, double Gz, R Rr, R h) {
    if (fabs(Gm) < kEps) return kPiS * Rr * Rr * h;
    if (fabs(Gp) < kEps && fabs(Gz) > kEps) return 2.0 * kPiS * Rr * Rr * xsin(0.5 * Gz * h) / Gz;
    if (fabs(Gz) < kEps && fabs(Gp) > kEps) return 2.0 * kPiS * Rr * h * xj1(Gp * Rr) / Gp;
    return 4.0 * kPiS * Rr * xsin(0.5 * Gz * h) * xj1(Gp * Rr) / (Gp * Gz);
}

// The rotated wave vector and the prefactor e^{-i G.O} / Omega (times the softening) of one shape.
struct Rotated {
    double Gx, Gy, Gz, Gm, Gp;
    c64 pref;
};
__device__ inline Rotated rotate(const ShapeDev &s, double Gx0, double Gy0, double Gz0) {
    const double *r = s.rot;
    Rotated o;
    o.Gx = r[0] * Gx0 + r[1] * Gy0 + r[2] * Gz0;
    o.Gy = r[3] * Gx0 + r[4] * Gy0 + r[5] * Gz0;
    o.Gz = r[6] * Gx0 + r[7] * Gy0 + r[8] * Gz0;
    o.Gm = sqrt(o.Gx * o.Gx + o.Gy * o.Gy + o.Gz * o.Gz);
    o.Gp = sqrt(o.Gx * o.Gx + o.Gy * o.Gy);
    double scale = s.inv_vol;
    if (s.soft != 0.0) scale *= exp(-s.soft * (o.Gx * o.Gx + o.Gy * o.Gy + o.Gz * o.Gz));
    o.pref = scale * cis_m(Gx0 * s.O[0] + Gy0 * s.O[1] + Gz0 * s.O[2]);
    return o;
}

// Shapes with closed-form transforms (0, 1, 2, 5-10, 16-18); L: the lengths (dual: seeded).
template <class R> __device__ CR<R> closed_form(const ShapeDev &s, const Rotated &g, const R *L) {
    const double Gx = g.Gx, Gy = g.Gy, Gz = g.Gz, Gm = g.Gm, Gp = g.Gp;
    const R L1 = L[0], L2 = L[1], L3 = L[2], L4 = L[3];
    CR<R> f = cr(R{0.0});
    switch (s.kind) {
    case 0:
    case 17:
    case 18: {   // cuboid, quantum well, quantum wire: a box of the host-resolved lengths
        R lb[3];
        for (int a = 0; a < 3; ++a) lb[a] = s.lbmap[a] >= 0 ? L[s.lbmap[a]] : R{s.Lb[a]};
        f = cr(sinc_len(Gx, lb[0]) * sinc_len(Gy, lb[1]) * sinc_len(Gz, lb[2]));
        break;
    }
    case 1:
        f = cr(fabs(Gm) < kEps ? 4.0 * kPiS * L1 * L1 * L1 / 3.0 : 4.0 * kPiS * L1 * L1 * L1 * sphere_kernel(Gm * L1));
        break;
    case 2: {
        const R gg = xsqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2 + Gz * Gz * L3 * L3);
        f = cr(fabs(val(gg)) < kEps ? 4.0 * kPiS * L1 * L2 * L3 / 3.0 : 4.0 * kPiS * L1 * L2 * L3 * sphere_kernel(gg));
        break;
    }
    case 5:
        for (int k = 0; k < 4; ++k) f = f + pyramid_segment(Gx, Gy, Gz, k, L1, L2, L3, 0);
        break;
    case 6: {   // the full pyramid minus its top
        const R h_full = L4 * L1 / (L1 - L3);
        CR<R> top = cr(R{0.0});
        for (int k = 0; k < 4; ++k) f = f + pyramid_segment(Gx, Gy, Gz, k, L1, L2, h_full, 0);
        const R t1 = L3, t2 = L2 * L3 / L1, th = L4 * L3 / (L1 - L3);
        for (int k = 0; k < 4; ++k) top = top + pyramid_segment(Gx, Gy, Gz, k, t1, t2, th, 0);
        f = f - cis_m(Gz * L4) * top;
        break;
    }
    case 7:
        for (int k = 0; k < s.nside; ++k) f = f + pyramid_segment(Gx, Gy, Gz, k, L1, R{0.0}, L2, s.nside);
        break;
    case 8: {
        CR<R> top = cr(R{0.0});
        for (int k = 0; k < s.nside; ++k) f = f + pyramid_segment(Gx, Gy, Gz, k, L1, R{0.0}, L1 * L3 / (L1 - L2), s.nside);
        for (int k = 0; k < s.nside; ++k) top = top + pyramid_segment(Gx, Gy, Gz, k, L2, R{0.0}, L3 * L2 / (L1 - L2), s.nside);
        f = f - cis_m(Gz * L3) * top;
        break;
    }
    case 9:
        f = cr(cylinder(Gm, Gp, Gz, L1, L2));
        break;
    case 10:
        f = cr(cylinder(Gm, Gp, Gz, L1, L3) - cylinder(Gm, Gp, Gz, L2, L3));
        break;
    case 16:
        for (int k = 0; k < s.nside; ++k) f = f + prism_segment(Gx, Gy, Gz, k, s.nside, L1, L2);
        break;
    default:
        break;
    }
    return f;
}

// Shapes whose transforms need 1-D integrals of Bessel functions (3, 4, 11-15).
// kKind: the shape type, a compile-time constant so that each instance holds one kind's registers.
template <int kKind, class R> __device__ CR<R> quadrature_form(const ShapeDev &s, const Rotated &g, const R *L, const double *gl) {
    const double Gx = g.Gx, Gy = g.Gy, Gz = g.Gz, Gm = g.Gm, Gp = g.Gp;
    const R L1 = L[0], L2 = L[1], L3 = L[2], L4 = L[3];
    CR<R> f = cr(R{0.0});
    switch (kKind) {
    case 3: {   // hemisphere, base centred on O, radius L1
        if (fabs(Gm) < kEps) { f = cr(2.0 * kPiS * L1 * L1 * L1 / 3.0); break; }
        const R sph = 2.0 * kPiS * L1 * L1 * L1 * sphere_kernel(Gm * L1);
        if (fabs(Gp) < kEps && fabs(Gz) > kEps) {
            const R a = L1 * Gz;
            f = cr(sph) - ci(2.0 * kPiS * L1 * L1 / Gz) * cr(0.5 - (a * xsin(a) + xcos(a) - 1.0) / (a * a));
        } else if (fabs(Gz) < kEps && fabs(Gp) > kEps) {
            f = cr(sph);
        } else {
            f = cr(sph) - ci(2.0 * kPiS * L1 * xj1(Gp * L1) / (Gp * Gz)) +
                ci(2.0 * kPiS * L1 * L1 * int_J0(Gz * L1, Gp * L1, gl) / Gz);
        }
        break;
    }
    case 4: {   // lens (half ellipsoid), base centred on O, semi-axes L1 L2 L3
        const R gm = xsqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2 + Gz * Gz * L3 * L3);
        const R gp = xsqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2);
        const R V2 = 2.0 * kPiS * L1 * L2 * L3;
        if (fabs(val(gm)) < kEps) { f = cr(V2 / 3.0); break; }
        const R sph = V2 * sphere_kernel(gm);
        const R a = Gz * L3;
        if (fabs(val(gp)) < kEps) f = cr(sph) - ci(2.0 * kPiS * L1 * L2 / Gz) * cr(0.5 - (xcos(a) + a * xsin(a) - 1.0) / (a * a));
        else if (fabs(Gz) < kEps) f = cr(sph);
        else f = cr(sph) - ci(2.0 * kPiS * L1 * L2 / Gz) * cr(xj1(gp) / gp - int_J0(a, gp, gl));
        break;
    }
    case 11:
        f = cone(Gm, R{Gp}, Gz, L1, L1 * L1, L2, gl);
        break;
    case 12:
        f = trunc_cone(Gm, Gp, Gz, L1, L2, L3, gl);
        break;
    case 13:
        f = trunc_cone(Gm, Gp, Gz, L1, L2, L4, gl) - cis_m(0.5 * Gz * L4) * cr(cylinder(Gm, Gp, Gz, L3, L4));
        break;
    case 14: {
        const R gq = xsqrt((Gx * L1) * (Gx * L1) + (Gy * L2) * (Gy * L2));
        f = cone(Gm, gq, Gz, R{1.0}, L1 * L2, L3, gl);
        break;
    }
    case 15: {
        const R gq = xsqrt((Gx * L1) * (Gx * L1) + (Gy * L2) * (Gy * L2));
        f = trunc_econe(Gm, gq, Gz, L1, L2, L3 / L1, L4, gl);
        break;
    }
    default:
        break;
    }
    return f;
}

// chi of one shape at G (R = Dd: with its derivative along the lengths' seeds); kKind: -1 for the
// closed forms (a run-time switch on the type), else the quadrature shape's type (one kernel instance
// per kind keeps each free of the others' registers)
template <int kKind, class R>
__device__ CR<R> shape_chi(const ShapeDev &s, double Gx0, double Gy0, double Gz0, const double *gl,
                           const double *seed) {
    const Rotated g = rotate(s, Gx0, Gy0, Gz0);
    R L[4];
    for (int k = 0; k < 4; ++k) L[k] = seeded<R>(s.L[k], seed ? seed[k] : 0.0);
    CR<R> f;
    if constexpr (kKind < 0) f = closed_form(s, g, L);
    else f = quadrature_form<kKind>(s, g, L, gl);
    return lift<R>(g.pref) * f;
}

// out[slot[j] * n + i] = chi of shape j at grid point i of the axes gx x gy x gz (all shapes of one family)
template <int kKind>
__global__ __launch_bounds__(256) void k_shape_grid(const ShapeDev *__restrict__ sd, const int *__restrict__ slot,
                                                    const double *__restrict__ gx, const double *__restrict__ gy,
                                                    const double *__restrict__ gz, int n0, int n1, int n2,
                                                    const double *__restrict__ gl, c64 *__restrict__ out) {
    const ShapeDev s = sd[blockIdx.y];
    const long n = (long)n0 * n1 * n2;
    c64 *o = out + (long)slot[blockIdx.y] * n;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        o[i] = shape_chi<kKind, double>(s, gx[ix], gy[iy], gz[iz], gl, nullptr);
    }
}

// The lengths' derivative seeds of one parameter (1 for each length it moves).
struct Seeds {
    double d[4];
};

// out[i] = chi of s at G[3 i ..], or with deriv its derivative along the seeds
template <int kKind>
__global__ __launch_bounds__(256) void k_shape_points(ShapeDev s, const double *__restrict__ G, int n,
                                                      const double *__restrict__ gl, int deriv, Seeds seeds,
                                                      c64 *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (deriv) {
        const CR<Dd> z = shape_chi<kKind, Dd>(s, G[3 * i], G[3 * i + 1], G[3 * i + 2], gl, seeds.d);
        out[i] = mk(z.re.d, z.im.d);
    } else {
        out[i] = shape_chi<kKind, double>(s, G[3 * i], G[3 * i + 1], G[3 * i + 2], gl, nullptr);
    }
}

// partial[j * gridDim.x + blockIdx.x] = Re sum_i grad_j[i] d chi_j[i] / d theta over the block's share
// of the grid, shape j = blockIdx.y with its gradient grid at slot[j]: the contraction of the
// gradient of chi_j with the exact derivative along the seeds (never stored)
template <int kKind>
__global__ __launch_bounds__(256) void k_shape_deriv_dots(const ShapeDev *__restrict__ sd, const int *__restrict__ slot,
                                                          const double *__restrict__ gx, const double *__restrict__ gy,
                                                          const double *__restrict__ gz, int n0, int n1, int n2,
                                                          const double *__restrict__ gl, Seeds seeds,
                                                          const c64 *__restrict__ grad, double *__restrict__ partial) {
    const ShapeDev s = sd[blockIdx.y];
    const long n = (long)n0 * n1 * n2;
    const c64 *g = grad + (long)slot[blockIdx.y] * n;
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        const CR<Dd> z = shape_chi<kKind, Dd>(s, gx[ix], gy[iy], gz[iz], gl, seeds.d);
        acc += g[i].re * z.re.d - g[i].im * z.im.d;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

}  // namespace

bool shape_on_device(int kind) { return kind >= 0 && kind <= 18; }

bool shape_needs_quadrature(int kind) { return kind == 3 || kind == 4 || (kind >= 11 && kind <= 15); }

int shape_family(int kind) { return shape_needs_quadrature(kind) ? kind : -1; }

namespace {
// f(std::integral_constant<int, family>) for a family key of shape_family
template <class F> void by_family(int family, F &&f) {
    switch (family) {
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 11: f(std::integral_constant<int, 11>{}); break;
    case 12: f(std::integral_constant<int, 12>{}); break;
    case 13: f(std::integral_constant<int, 13>{}); break;
    case 14: f(std::integral_constant<int, 14>{}); break;
    case 15: f(std::integral_constant<int, 15>{}); break;
    default: f(std::integral_constant<int, -1>{}); break;
    }
}
}  // namespace

ShapeDev shape_device_params(const qd_inputs &in, int idx, Cell cell) {
    const qd_shape &s = in.shapes[idx];
    const qd_shape &c0 = in.shapes[0];
    if (!shape_on_device(s.shape)) throw std::logic_error("shape type " + std::to_string(s.shape) + " has no device form");
    ShapeDev d{};
    d.kind = s.shape;
    d.O[0] = s.Ox; d.O[1] = s.Oy; d.O[2] = s.Oz;
    d.L[0] = s.L1; d.L[1] = s.L2; d.L[2] = s.L3; d.L[3] = s.L4;
    double vol;
    if (cell == Cell::Coulomb) vol = in.CL[0] * in.CL[1] * in.CL[2];
    else if (cell == Cell::Electronic || idx == 0) vol = c0.L1 * c0.L2 * c0.L3;
    else vol = in.SL[0] * in.SL[1] * in.SL[2];
    d.inv_vol = 1.0 / vol;
    d.soft = in.softening ? 0.5 * in.sigma_softening * in.sigma_softening : 0.0;
    // proper Euler rotation z, y', z'' (degrees)
    const double sa = std::sin(kPiS * s.alpha / 180), ca = std::cos(kPiS * s.alpha / 180);
    const double sb = std::sin(kPiS * s.beta / 180), cb = std::cos(kPiS * s.beta / 180);
    const double sg = std::sin(kPiS * s.gamma / 180), cg = std::cos(kPiS * s.gamma / 180);
    const double rot[9] = {cb * ca * cg - sa * sg,  cg * cb * sa + sg * ca,  -sb * cg,
                           -cb * ca * sg - sa * cg, -sg * cb * sa + cg * ca, sb * sg,
                           sb * ca,                 sa * sb,                 cb};
    for (int k = 0; k < 9; ++k) d.rot[k] = rot[k];
    if (s.shape == 7 || s.shape == 8) {
        d.nside = (int)std::lround(0.1 * (s.shape == 7 ? s.L3 : s.L4));
        if (d.nside < 3) throw std::runtime_error("pyramid needs at least 3 sides");
    }
    if (s.shape == 16) d.nside = (int)std::lround(0.1 * s.L3);
    d.lbmap[0] = d.lbmap[1] = d.lbmap[2] = -1;
    if (s.shape == 0) {
        d.Lb[0] = s.L1; d.Lb[1] = s.L2; d.Lb[2] = s.L3;
        d.lbmap[0] = 0; d.lbmap[1] = 1; d.lbmap[2] = 2;
    }
    if (s.shape == 17 || s.shape == 18) {
        const int dir = (int)std::lround(0.1 * s.L1);
        if (cell == Cell::Electronic) { d.Lb[0] = c0.L1; d.Lb[1] = c0.L2; d.Lb[2] = c0.L3; }
        else { d.Lb[0] = in.SL[0]; d.Lb[1] = in.SL[1]; d.Lb[2] = in.SL[2]; }
        if (dir < 0 || dir > 2)
            throw std::runtime_error(std::string("quantum-") + (s.shape == 17 ? "well" : "wire") +
                                     " direction must be 0, 1 or 2");
        if (s.shape == 17) {
            d.Lb[dir] = s.L2;
            d.lbmap[dir] = 1;
        } else {
            const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
            d.Lb[o1] = s.L2;
            d.Lb[o2] = s.L3;
            d.lbmap[o1] = 1;
            d.lbmap[o2] = 2;
        }
    }
    return d;
}

void launch_shape_grid(const ShapeDev *sd, const int *slot, int count, int family, const double *const g[3],
                       const int n[3], const double *gl, c64 *out, hipStream_t s) {
    if (count == 0) return;
    const long np = (long)n[0] * n[1] * n[2];
    const unsigned gx = (unsigned)std::max<long>(1, std::min<long>(4096, (np + 255) / 256));
    by_family(family, [&](auto K) {
        hipLaunchKernelGGL(k_shape_grid<decltype(K)::value>, dim3(gx, (unsigned)count), dim3(256), 0, s, sd, slot, g[0],
                           g[1], g[2], n[0], n[1], n[2], gl, out);
    });
    QD_LAUNCH_CHECK();
}

bool length_seeds(int param, int kind, double seed[4]) {
    for (int k = 0; k < 4; ++k) seed[k] = 0.0;
    switch (param) {
    case QD_DP_L1: seed[0] = 1.0; break;
    case QD_DP_L2: seed[1] = 1.0; break;
    case QD_DP_L3: seed[2] = 1.0; break;
    case QD_DP_L4: seed[3] = 1.0; break;
    case QD_DP_L12: seed[0] = seed[1] = 1.0; break;
    default: return false;
    }
    // the quantum well's and wire's direction is selected by Length 1: it is not a length there
    if (kind == 17 || kind == 18) seed[0] = 0.0;
    return true;
}

void launch_shape_points(const ShapeDev &sd, const double *G, int n, const double *gl, const double *seed, c64 *out,
                         hipStream_t s) {
    if (n == 0) return;
    const dim3 grid((unsigned)((n + 255) / 256));
    Seeds sv{{0.0, 0.0, 0.0, 0.0}};
    if (seed)
        for (int k = 0; k < 4; ++k) sv.d[k] = seed[k];
    const int deriv = seed ? 1 : 0;
    by_family(shape_family(sd.kind), [&](auto K) {
        hipLaunchKernelGGL(k_shape_points<decltype(K)::value>, grid, dim3(256), 0, s, sd, G, n, gl, deriv, sv, out);
    });
    QD_LAUNCH_CHECK();
}


void launch_shape_deriv_dots(const ShapeDev *sd, const int *slot, int count, int family, const double *const g[3],
                             const int n[3], const double *gl, const double seed[4], const c64 *grad, double *partial,
                             int nblk, hipStream_t s) {
    if (count == 0) return;
    Seeds sv{{seed[0], seed[1], seed[2], seed[3]}};
    by_family(family, [&](auto K) {
        hipLaunchKernelGGL(k_shape_deriv_dots<decltype(K)::value>, dim3((unsigned)nblk, (unsigned)count), dim3(256), 0, s,
                           sd, slot, g[0], g[1], g[2], n[0], n[1], n[2], gl, sv, grad, partial);
    });
    QD_LAUNCH_CHECK();
}

}  // namespace qd
