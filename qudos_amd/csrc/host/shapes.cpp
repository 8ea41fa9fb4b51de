// Fourier transforms of shape characteristic functions, chi_s(G) = (1/Omega) int_s e^{-iG.r}.
// Restates characteristic_function_fourier.c:58-1105 (prefactor, Euler rotation, softening,
// per-shape closed forms and their G -> 0 limits) and the 1-D integrals of
// characteristic_function_fourier_integrals.c.  The reference evaluates those integrals with
// GSL QAG (61-point Gauss-Kronrod, tolerance 1e-4); every integrand here is analytic on its
// interval, so this file uses its own Gauss-Legendre rule with an interval-halving error control
// at ~1e-14, which reproduces the converged integral (see the quadrature parity tests).
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "qd_dual.hpp"
#include "qd_host.hpp"

namespace qd {
namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- Gauss-Legendre quadrature
struct GLRule {
    std::vector<double> x, w;   // nodes / weights on [-1, 1]
};

GLRule make_gl(int n) {
    GLRule r;
    r.x.resize(n);
    r.w.resize(n);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; ++k) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        r.x[i] = -z;
        r.x[n - 1 - i] = z;
        r.w[i] = r.w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    return r;
}

const GLRule &gl_lo() { static const GLRule r = make_gl(40); return r; }
const GLRule &gl_hi() { static const GLRule r = make_gl(64); return r; }
}  // namespace

std::vector<double> gauss_legendre_rules() {
    std::vector<double> r;
    for (const GLRule *g : {&gl_lo(), &gl_hi()}) {
        r.insert(r.end(), g->x.begin(), g->x.end());
        r.insert(r.end(), g->w.begin(), g->w.end());
    }
    return r;
}

namespace {

// Real scalar R (double, or Dual for derivatives with respect to one shape parameter) and its
// complex partner C.  Helpers keep the double instantiation's operations as before.
template <class R> using CofR = typename CplxOf<R>::type;
inline double bj0(double x) { return ::j0(x); }
inline double bj1(double x) { return ::j1(x); }
inline Dual bj0(Dual x) { return j0(x); }
inline Dual bj1(Dual x) { return j1(x); }
inline cplx mkc(double re, double im) { return cplx(re, im); }
inline CDual mkc(Dual re, Dual im) { return CDual(re, im); }
using std::cos;
using std::sin;
using std::sqrt;
using std::atan;

template <class R, class F>
CofR<R> gl_apply(const GLRule &r, const F &f, R a, R b) {
    const R h = 0.5 * (b - a), c = 0.5 * (b + a);
    CofR<R> s = 0.0;
    for (size_t i = 0; i < r.x.size(); ++i) s += r.w[i] * f(c + h * r.x[i]);
    return h * s;
}

template <class R, class F>
CofR<R> integrate(const F &f, R a, R b, int depth = 0) {
    const CofR<R> lo = gl_apply<R>(gl_lo(), f, a, b);
    const CofR<R> hi = gl_apply<R>(gl_hi(), f, a, b);
    if (std::abs(val(hi) - val(lo)) <= 1e-14 * std::max(1.0, std::abs(val(hi))) || depth >= 12) return hi;
    const R m = 0.5 * (a + b);
    return integrate<R>(f, a, m, depth + 1) + integrate<R>(f, m, b, depth + 1);
}

template <class R> R re_part(const CofR<R> &z);
template <> double re_part<double>(const cplx &z) { return z.real(); }
template <> Dual re_part<Dual>(const CDual &z) { return Dual(z.v.real(), z.d.real()); }

// int_0^1 x cos(alpha sqrt(1-x^2)) J0(beta x) dx
template <class R>
R int_J0(R alpha, R beta) {
    auto f = [&](R x) {
        const R u = 1.0 - x * x;
        return CofR<R>(x * cos(alpha * sqrt(val(u) > 0 ? u : R(0.0))) * bj0(beta * x));
    };
    return re_part<R>(integrate<R>(f, R(0.0), R(1.0)));
}

// int_{R1/(R2-R1)}^{R2/(R2-R1)} x J0((R2-R1) alpha x) e^{i beta x} dx
template <class R>
CofR<R> int_TC(R alpha, R R1, R R2, R beta) {
    auto f = [&](R x) { return (x * bj0((R2 - R1) * alpha * x)) * mkc(cos(beta * x), sin(beta * x)); };
    return integrate<R>(f, R1 / (R2 - R1), R2 / (R2 - R1));
}

// int over the same limits of x^2 J0((R2-R1) alpha x)
template <class R>
R int_TC_z(R alpha, R R1, R R2) {
    auto f = [&](R x) { return CofR<R>(x * x * bj0((R2 - R1) * alpha * x)); };
    return re_part<R>(integrate<R>(f, R1 / (R2 - R1), R2 / (R2 - R1)));
}

// int_0^1 x J0(alpha x) e^{i beta x} dx (cone, elliptical cone)
template <class R>
CofR<R> int_2(R alpha, R beta) {
    auto f = [&](R x) { return (x * bj0(alpha * x)) * mkc(cos(beta * x), sin(beta * x)); };
    return integrate<R>(f, R(0.0), R(1.0));
}

// int_0^1 x^2 J0(alpha x) dx
template <class R>
R int_2_z(R alpha) {
    auto f = [&](R x) { return CofR<R>(x * x * bj0(alpha * x)); };
    return re_part<R>(integrate<R>(f, R(0.0), R(1.0)));
}

// int_f^1 x J0(alpha x) e^{i beta x / (1 - f)} dx (truncated elliptical cone)
template <class R>
CofR<R> int_2_etc(R alpha, R beta, R fr) {
    auto f = [&](R x) {
        const R ph = beta * x / (1.0 - fr);
        return (x * bj0(alpha * x)) * mkc(cos(ph), sin(ph));
    };
    return integrate<R>(f, fr, R(1.0));
}

// int_f^1 x^2 J0(alpha x) dx
template <class R>
R int_2_etc_z(R alpha, R fr) {
    auto f = [&](R x) { return CofR<R>(x * x * bj0(alpha * x)); };
    return re_part<R>(integrate<R>(f, fr, R(1.0)));
}

template <class R> inline CofR<R> cis_m(R a) { return mkc(cos(a), -sin(a)); }   // e^{-ia}
template <class R> inline CofR<R> expi(R a) { return mkc(cos(a), sin(a)); }      // e^{ia}
template <class R> inline CofR<R> imag_unit(R x) { return mkc(R(0.0), x); }       // i x
inline cplx times_i(const cplx &z) { return z * cplx(0.0, 1.0); }
inline CDual times_i(const CDual &z) { return CDual(z.v * cplx(0.0, 1.0), z.d * cplx(0.0, 1.0)); }
template <class R> inline R sinc_len(double G, R L) {   // L sin(GL/2)/(GL/2)
    const R t = 0.5 * G * L;
    return val(t) == 0.0 ? L : L * sin(t) / t;
}
const cplx I1(0.0, 1.0);

// (e^{iGa} - 1)/(iG), a at G = 0 (characteristic_function_fourier_integrals.c:28-50)
template <class R>
CofR<R> edge_integral(R a, R G) {
    if (std::fabs(val(G)) < kEps) return CofR<R>(a);
    return (expi<R>(G * a) - 1.0) / imag_unit<R>(G);
}

// Triangular pyramid segment (apex height h over the base plane, base edge at x0 with half-width
// y0), side k (characteristic_function_fourier.c:1112-1271).  nside == 0: rectangular base of half
// lengths L1, L2, the segment turned by 2 atan(L2/L1) k or 2 atan(L1/L2) k as the reference does
// (a tiling of the base for square bases); nside > 0: regular polygon of apothem L1, turned by
// 2 pi k / nside.
template <class R>
CofR<R> pyramid_segment(double Gx0, double Gy0, double Gz, int k, R L1, R L2, R h, int nside = 0) {
    using C = CofR<R>;
    R alpha, x0, y0;
    if (nside > 0) { alpha = R(kPi / nside); x0 = L1; y0 = std::tan(kPi / nside) * L1; }
    else if (k == 0 || k == 2) { alpha = atan(L2 / L1); x0 = L1; y0 = L2; }
    else { alpha = atan(L1 / L2); x0 = L2; y0 = L1; }
    const R ang = 2.0 * alpha * (double)k;
    const R Gx = cos(ang) * Gx0 + sin(ang) * Gy0;
    const R Gy = -sin(ang) * Gx0 + cos(ang) * Gy0;
    const R Gm = sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const bool zx = std::fabs(val(Gx)) < kEps, zy = std::fabs(val(Gy)) < kEps, zz = std::fabs(Gz) < kEps;
    if (val(Gm) < kEps) return C(x0 * y0 * h / 3.0);
    if (zx && zy)
        return times_i(C(2.0 * y0 * x0 / Gz)) *
               ((1.0 - imag_unit<R>(Gz * h) - cis_m<R>(Gz * h)) / (Gz * Gz * h * h) - 0.5);
    if (zx && zz) return C((2.0 * h * x0 / (Gy * Gy * y0)) * (1.0 - sin(Gy * y0) / (Gy * y0)));
    if (zy && zz)
        return (2.0 * h * y0 / (Gx * Gx * x0)) *
               (-1.0 - cis_m<R>(Gx * x0) - imag_unit<R>(2.0 / (Gx * x0)) + times_i(2.0 * cis_m<R>(Gx * x0) / (Gx * x0)));
    if (zx) {
        const R sp = Gy * y0 + Gz * h, sm = Gy * y0 - Gz * h;
        const C pre = imag_unit<R>(-x0 / (Gz * Gy));
        const C tail = C(-2.0 * (cos(Gy * y0) - 1.0) / (Gy * y0));
        C body;
        if (std::fabs(val(sm)) < kEps) body = C(cplx(0.0, -1.0)) + (expi<R>(sp) - 1.0) / sp;
        else if (std::fabs(val(sp)) < kEps) body = (cis_m<R>(sm) - 1.0) / sm + C(I1);
        else body = (cis_m<R>(sm) - 1.0) / sm + (expi<R>(sp) - 1.0) / sp;
        return pre * (cis_m<R>(Gz * h) * body + tail);
    }
    if (zy) {
        const R q = Gx - Gz * h / x0;
        const C base = (cis_m<R>(Gx * x0) * (1.0 + imag_unit<R>(Gx * x0)) - 1.0) / (Gx * Gx);
        const C pre = imag_unit<R>(2.0 * y0 / (Gz * x0));
        if (std::fabs(val(q)) < kEps) return pre * (0.5 * cis_m<R>(Gz * h) * (x0 * x0) - base);
        const R r = Gx * x0 - Gz * h;
        return pre * (cis_m<R>(Gz * h) * (cis_m<R>(r) * (1.0 + imag_unit<R>(r)) - 1.0) / (q * q) - base);
    }
    if (zz) {
        const R pp = Gx + Gy * y0 / x0, pm = Gx - Gy * y0 / x0;
        const C Ap = imag_unit<R>(1.0 / (x0 * pp * pp)) * (cis_m<R>(x0 * pp) - 1.0);
        const C Am = imag_unit<R>(1.0 / (x0 * pm * pm)) * (cis_m<R>(x0 * pm) - 1.0);
        if (std::fabs(val(pm)) < kEps) return (h / Gy) * (1.0 / pp - Ap - imag_unit<R>(0.5 * x0));
        if (std::fabs(val(pp)) < kEps) return (h / Gy) * (-1.0 / pm + Am + imag_unit<R>(0.5 * x0));
        return (h / Gy) * (1.0 / pp - 1.0 / pm - Ap + Am);
    }
    const R s1 = Gy * y0 / x0, s2 = Gz * h / x0;
    return (-1.0 / (Gy * Gz)) * (edge_integral<R>(x0, -Gx - s1 + s2) * cis_m<R>(Gz * h) -
                                 edge_integral<R>(x0, -Gx + s1 + s2) * cis_m<R>(Gz * h) -
                                 edge_integral<R>(x0, -Gx - s1) + edge_integral<R>(x0, -Gx + s1));
}

// One of N wedge segments of a regular N-sided prism of inradius x0 and height h centred in z
// (characteristic_function_fourier.c:1276-1396).
template <class R>
CofR<R> prism_segment(double Gx0, double Gy0, double Gz, int k, int N, R x0, R h) {
    using C = CofR<R>;
    const double alpha = kPi / N;
    const double ang = 2.0 * kPi * k / N;
    const double Gx = std::cos(ang) * Gx0 + std::sin(ang) * Gy0;
    const double Gy = -std::sin(ang) * Gx0 + std::cos(ang) * Gy0;
    const double Gm = std::sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const double ta = std::tan(alpha);
    const bool zx = std::fabs(Gx) < kEps, zy = std::fabs(Gy) < kEps, zz = std::fabs(Gz) < kEps;
    if (Gm < kEps) return C(x0 * x0 * h * ta);
    if (zx && zy) return C(2.0 * x0 * x0 * ta * sin(0.5 * Gz * h) / Gz);
    if (zx && zz) return C((2.0 * h / (Gy * Gy * ta)) * (1.0 - cos(Gy * x0 * ta)));
    const C xterm = imag_unit<R>(x0) * cis_m<R>(Gx * x0) + (1.0 / Gx) * (cis_m<R>(Gx * x0) - 1.0);
    if (zy && zz) return (2.0 * h * ta / Gx) * xterm;
    if (zx) return C((4.0 * sin(0.5 * Gz * h) / (Gy * Gy * Gz * ta)) * (1.0 - cos(Gy * x0 * ta)));
    if (zy) return (4.0 * sin(0.5 * Gz * h) * ta / (Gz * Gx)) * xterm;
    const double up = Gy * ta + Gx, um = -Gy * ta + Gx;
    const C ep = 1.0 - cis_m<R>(Gy * x0 * ta + Gx * x0);
    const C em = 1.0 - cis_m<R>(-Gy * x0 * ta + Gx * x0);
    C face;
    if (std::fabs(up) < kEps) face = imag_unit<R>(x0) - em / um;
    else if (std::fabs(um) < kEps) face = ep / up - imag_unit<R>(x0);
    else face = ep / up - em / um;
    if (zz) return (h / Gy) * face;
    return ((2.0 / (Gy * Gz)) * sin(0.5 * Gz * h)) * face;
}

// Cylinder of radius Rr and height h centred on its origin (characteristic_function_fourier.c:556-588).
template <class R>
CofR<R> cylinder_t(double Gm, double Gp, double Gz, R Rr, R h) {
    using C = CofR<R>;
    if (std::fabs(Gm) < kEps) return C(kPi * Rr * Rr * h);
    if (std::fabs(Gp) < kEps && std::fabs(Gz) > kEps) return C(2.0 * kPi * Rr * Rr * sin(0.5 * Gz * h) / Gz);
    if (std::fabs(Gz) < kEps && std::fabs(Gp) > kEps) return C(2.0 * kPi * Rr * h * bj1(Gp * Rr) / Gp);
    return C(4.0 * kPi * Rr * sin(0.5 * Gz * h) * bj1(Gp * Rr) / (Gp * Gz));
}

// Cone over a disc (shape 11: gq = |G_par|, Rr = radius, A = B = Rr^2) or an ellipse (shape 14:
// gq = |(Gx L1, Gy L2)|, Rr = 1, A = B = L1 L2), height H, base on the origin
// (characteristic_function_fourier.c:622-667, 769-823); a = gq Rr is the Bessel argument.
template <class R>
CofR<R> cone_t(double Gm, R gq, double Gz, R Rr, R A, R H) {
    using C = CofR<R>;
    if (std::fabs(Gm) < kEps) return C(kPi * A * H / 3.0);
    if (std::fabs(val(gq)) < kEps)
        return imag_unit<R>(2.0 * kPi * A / Gz) * ((1.0 - imag_unit<R>(Gz * H) - cis_m<R>(Gz * H)) / ((Gz * H) * (Gz * H)) - 0.5);
    const R a = gq * Rr;
    if (std::fabs(Gz) < kEps) return C(2.0 * kPi * A * H * (bj1(a) / a - int_2_z<R>(a)));
    return imag_unit<R>(2.0 * kPi * A / Gz) * (cis_m<R>(Gz * H) * int_2<R>(a, R(Gz * H))) -
           imag_unit<R>(2.0 * kPi * A * bj1(a) / (a * Gz));
}

// Truncated cone: base radius R1 on the origin, top radius R2 at height H
// (characteristic_function_fourier.c:671-718).
template <class R>
CofR<R> trunc_cone_t(double Gm, double Gp, double Gz, R R1, R R2, R H) {
    using C = CofR<R>;
    const R dR = R1 - R2;
    if (std::fabs(Gm) < kEps) return C(kPi * H * (R1 * R1 + R1 * R2 + R2 * R2) / 3.0);
    if (std::fabs(Gp) < kEps) {
        const R a = Gz * H;
        return imag_unit<R>(2.0 * kPi / Gz) *
               (0.5 * (R2 * R2) * (cis_m<R>(a) - 1.0) +
                ((dR * dR) / (a * a)) * (cis_m<R>(a) * (imag_unit<R>(Gz * R2 * H / dR) - 1.0) +
                                         (1.0 - imag_unit<R>(R1 * Gz * H / dR))) -
                0.5 * (R1 * R1 - R2 * R2));
    }
    const R jj = (R1 * bj1(Gp * R1) - R2 * bj1(Gp * R2)) / Gp;
    if (std::fabs(Gz) < kEps)
        return C((2.0 * kPi * H) * (R2 * bj1(Gp * R2) / Gp - dR * dR * int_TC_z<R>(R(Gp), R2, R1) + (R1 / dR) * jj));
    return imag_unit<R>(2.0 * kPi / Gz) * (R2 * bj1(Gp * R2) / Gp * (cis_m<R>(Gz * H) - 1.0) +
                                         (dR * dR) * cis_m<R>(Gz * H * R1 / dR) * int_TC<R>(R(Gp), R2, R1, R(Gz * H)) -
                                         jj);
}

// Truncated elliptical cone: base semi-axes L1, L2, top L1 fr, L2 fr at height H; gq = |(Gx L1,
// Gy L2)| (characteristic_function_fourier.c:825-884).
template <class R>
CofR<R> trunc_econe_t(double Gm, R gq, double Gz, R L1, R L2, R fr, R H) {
    using C = CofR<R>;
    const R A = L1 * L2;
    if (std::fabs(Gm) < kEps) return C(kPi * A * H * (1.0 + fr + fr * fr) / 3.0);
    const C e = cis_m<R>(Gz * H);
    if (std::fabs(val(gq)) < kEps)
        return imag_unit<R>(2.0 * kPi * A / Gz) *
               (0.5 * fr * fr * (e - 1.0) +
                ((fr - 1.0) / (Gz * Gz * H * H)) * (e * (-imag_unit<R>(Gz * H * fr) - fr + 1.0) + (imag_unit<R>(Gz * H) + fr - 1.0)) -
                0.5 * (1.0 - fr * fr));
    const R jb = (bj1(gq) - fr * bj1(fr * gq)) / gq;
    if (std::fabs(Gz) < kEps)
        return C(2.0 * kPi * A * H * (fr * bj1(fr * gq) / gq - int_2_etc_z<R>(gq, fr) / (1.0 - fr) + jb / (1.0 - fr)));
    return imag_unit<R>(2.0 * kPi * A / Gz) *
           (fr * (e - 1.0) * (bj1(fr * gq) / gq) + cis_m<R>(Gz * H / (1.0 - fr)) * int_2_etc<R>(gq, R(Gz * H), fr) - jb);
}

// 4 pi (sin g - g cos g)/g^3 scaled: the unit-sphere transform.
template <class R> inline R sphere_kernel(R g) { return (sin(g) - g * cos(g)) / (g * g * g); }

// Geometry of one shape that derivatives may vary (centre and lengths).
template <class R> struct ShapeVar { R Ox, Oy, Oz, L1, L2, L3, L4; };

template <class R>
CofR<R> shape_transform_t(const qd_inputs &in, int idx, double Gx0, double Gy0, double Gz0, Cell cell,
                          const ShapeVar<R> &g) {
    using C = CofR<R>;
    const qd_shape &s = in.shapes[idx];
    const qd_shape &c0 = in.shapes[0];
    double vol;
    if (cell == Cell::Coulomb) vol = in.CL[0] * in.CL[1] * in.CL[2];
    else if (cell == Cell::Electronic || idx == 0) vol = c0.L1 * c0.L2 * c0.L3;
    else vol = in.SL[0] * in.SL[1] * in.SL[2];
    C pref = cis_m<R>(Gx0 * g.Ox + Gy0 * g.Oy + Gz0 * g.Oz) / vol;

    // proper Euler rotation z, y', z'' (angles in degrees)
    const double sa = std::sin(kPi * s.alpha / 180), ca = std::cos(kPi * s.alpha / 180);
    const double sb = std::sin(kPi * s.beta / 180), cb = std::cos(kPi * s.beta / 180);
    const double sg = std::sin(kPi * s.gamma / 180), cg = std::cos(kPi * s.gamma / 180);
    const double Gx = (cb * ca * cg - sa * sg) * Gx0 + (cg * cb * sa + sg * ca) * Gy0 + (-sb * cg) * Gz0;
    const double Gy = (-cb * ca * sg - sa * cg) * Gx0 + (-sg * cb * sa + cg * ca) * Gy0 + (sb * sg) * Gz0;
    const double Gz = (sb * ca) * Gx0 + (sa * sb) * Gy0 + cb * Gz0;
    const double Gm = std::sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const double Gp = std::sqrt(Gx * Gx + Gy * Gy);
    if (in.softening)
        pref = pref * std::exp(-0.5 * in.sigma_softening * in.sigma_softening * (Gx * Gx + Gy * Gy + Gz * Gz));

    const R L1 = g.L1, L2 = g.L2, L3 = g.L3, L4 = g.L4;
    C f;
    switch (s.shape) {
    case 0:   // cuboid centred on O
        f = C(sinc_len<R>(Gx, L1) * sinc_len<R>(Gy, L2) * sinc_len<R>(Gz, L3));
        break;
    case 1:   // sphere of radius L1
        f = C(std::fabs(Gm) < kEps ? 4.0 * kPi * L1 * L1 * L1 / 3.0
                                    : 4.0 * kPi * L1 * L1 * L1 * sphere_kernel<R>(Gm * L1));
        break;
    case 2: { // ellipsoid, semi-axes L1 L2 L3
        const R gg = sqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2 + Gz * Gz * L3 * L3);
        f = C(std::fabs(val(gg)) < kEps ? 4.0 * kPi * L1 * L2 * L3 / 3.0
                                        : 4.0 * kPi * L1 * L2 * L3 * sphere_kernel<R>(gg));
        break;
    }
    case 3: { // hemisphere, base centred on O, radius L1
        const R Rr = L1;
        const R sph = 2.0 * kPi * Rr * Rr * Rr * sphere_kernel<R>(Gm * Rr);
        if (std::fabs(Gm) < kEps) f = C(2.0 * kPi * Rr * Rr * Rr / 3.0);
        else if (std::fabs(Gp) < kEps && std::fabs(Gz) > kEps) {
            const R a = Rr * Gz;
            f = sph - imag_unit<R>(2.0 * kPi * Rr * Rr / Gz) * (0.5 - (a * sin(a) + cos(a) - 1.0) / (a * a));
        } else if (std::fabs(Gz) < kEps && std::fabs(Gp) > kEps) f = C(sph);
        else
            f = sph - imag_unit<R>((2.0 * kPi * Rr * bj1(Gp * Rr)) / (Gp * Gz)) +
                imag_unit<R>((2.0 * kPi * Rr * Rr * int_J0<R>(Gz * Rr, Gp * Rr)) / Gz);
        break;
    }
    case 4: { // lens (half ellipsoid), base centred on O, semi-axes L1 L2 L3
        const R gm = sqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2 + Gz * Gz * L3 * L3);
        const R gp = sqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2);
        const R V2 = 2.0 * kPi * L1 * L2 * L3;
        if (std::fabs(val(gm)) < kEps) { f = C(V2 / 3.0); break; }
        const R sph = V2 * sphere_kernel<R>(gm);
        const R a = Gz * L3;
        if (std::fabs(val(gp)) < kEps)
            f = sph - imag_unit<R>(2.0 * kPi * L1 * L2 / Gz) * (0.5 - (cos(a) + a * sin(a) - 1.0) / (a * a));
        else if (std::fabs(Gz) < kEps) f = C(sph);
        else f = sph - imag_unit<R>(2.0 * kPi * L1 * L2 / Gz) * (bj1(gp) / gp - int_J0<R>(a, gp));
        break;
    }
    case 6: { // truncated rectangular pyramid = full pyramid minus its top
        const R h_full = L4 * L1 / (L1 - L3);
        C acc = 0.0;
        for (int k = 0; k < 4; ++k) acc += pyramid_segment<R>(Gx, Gy, Gz, k, L1, L2, h_full);
        const R t1 = L3, t2 = L2 * L3 / L1, th = L4 * L3 / (L1 - L3);
        C top = 0.0;
        for (int k = 0; k < 4; ++k) top += pyramid_segment<R>(Gx, Gy, Gz, k, t1, t2, th);
        f = acc - cis_m<R>(Gz * L4) * top;
        break;
    }
    case 12:  // truncated cone: base radius L1 at O, top radius L2 at height L3
        f = trunc_cone_t<R>(Gm, Gp, Gz, L1, L2, L3);
        break;
    case 5: { // rectangular-base pyramid: half base lengths L1, L2, height L3
        C acc = 0.0;
        for (int k = 0; k < 4; ++k) acc += pyramid_segment<R>(Gx, Gy, Gz, k, L1, L2, L3);
        f = acc;
        break;
    }
    case 7: { // regular pyramid: apothem L1, height L2, L3/10 sides
        const int nside = (int)std::lround(0.1 * val(L3));
        if (nside < 3) throw std::runtime_error("pyramid needs at least 3 sides");
        C acc = 0.0;
        for (int k = 0; k < nside; ++k) acc += pyramid_segment<R>(Gx, Gy, Gz, k, L1, R(0.0), L2, nside);
        f = acc;
        break;
    }
    case 8: { // regular truncated pyramid: apothems L1 (base), L2 (top), height L3, L4/10 sides
        const int nside = (int)std::lround(0.1 * val(L4));
        if (nside < 3) throw std::runtime_error("pyramid needs at least 3 sides");
        C acc = 0.0, top = 0.0;
        for (int k = 0; k < nside; ++k) acc += pyramid_segment<R>(Gx, Gy, Gz, k, L1, R(0.0), L1 * L3 / (L1 - L2), nside);
        for (int k = 0; k < nside; ++k) top += pyramid_segment<R>(Gx, Gy, Gz, k, L2, R(0.0), L3 * L2 / (L1 - L2), nside);
        f = acc - cis_m<R>(Gz * L3) * top;
        break;
    }
    case 9:   // cylinder: radius L1, height L2, centred on O
        f = cylinder_t<R>(Gm, Gp, Gz, L1, L2);
        break;
    case 10:  // cylindrical ring: radii L1 (outer), L2 (inner), height L3, centred on O
        f = cylinder_t<R>(Gm, Gp, Gz, L1, L3) - cylinder_t<R>(Gm, Gp, Gz, L2, L3);
        break;
    case 11:  // cone: base radius L1 at O, height L2
        f = cone_t<R>(Gm, R(Gp), Gz, L1, L1 * L1, L2);
        break;
    case 13:  // truncated conic ring: truncated cone (L1, L2, height L4) less a cylinder of radius L3
        f = trunc_cone_t<R>(Gm, Gp, Gz, L1, L2, L4) - cis_m<R>(0.5 * Gz * L4) * cylinder_t<R>(Gm, Gp, Gz, L3, L4);
        break;
    case 14: { // elliptical cone: base semi-axes L1, L2 at O, height L3
        const R gq = sqrt((Gx * L1) * (Gx * L1) + (Gy * L2) * (Gy * L2));
        f = cone_t<R>(Gm, gq, Gz, R(1.0), L1 * L2, L3);
        break;
    }
    case 15: { // truncated elliptical cone: base semi-axes L1, L2, top semi-axis L3 along x, height L4
        const R gq = sqrt((Gx * L1) * (Gx * L1) + (Gy * L2) * (Gy * L2));
        f = trunc_econe_t<R>(Gm, gq, Gz, L1, L2, L3 / L1, L4);
        break;
    }
    case 16: { // regular prism: inradius L1, height L2, L3/10 sides
        const int nside = (int)std::lround(0.1 * val(L3));
        C acc = 0.0;
        for (int k = 0; k < nside; ++k) acc += prism_segment<R>(Gx, Gy, Gz, k, nside, L1, L2);
        f = acc;
        break;
    }
    case 17: { // quantum well along axis lround(L1/10), thickness L2, spanning the cell otherwise
        const int d = (int)std::lround(0.1 * val(L1));
        R Ls[3];
        if (cell == Cell::Electronic) { Ls[0] = c0.L1; Ls[1] = c0.L2; Ls[2] = c0.L3; }
        else { Ls[0] = in.SL[0]; Ls[1] = in.SL[1]; Ls[2] = in.SL[2]; }
        if (d < 0 || d > 2) throw std::runtime_error("quantum-well direction must be 0, 1 or 2");
        Ls[d] = L2;
        f = C(sinc_len<R>(Gx, Ls[0]) * sinc_len<R>(Gy, Ls[1]) * sinc_len<R>(Gz, Ls[2]));
        break;
    }
    case 18: { // quantum wire along axis lround(L1/10), cross-section L2 x L3 over the other two axes
        const int d = (int)std::lround(0.1 * val(L1));
        R Ls[3];
        if (cell == Cell::Electronic) { Ls[0] = c0.L1; Ls[1] = c0.L2; Ls[2] = c0.L3; }
        else { Ls[0] = in.SL[0]; Ls[1] = in.SL[1]; Ls[2] = in.SL[2]; }
        if (d < 0 || d > 2) throw std::runtime_error("quantum-wire direction must be 0, 1 or 2");
        const int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
        Ls[o1] = L2;
        Ls[o2] = L3;
        f = C(sinc_len<R>(Gx, Ls[0]) * sinc_len<R>(Gy, Ls[1]) * sinc_len<R>(Gz, Ls[2]));
        break;
    }
    default:
        throw std::runtime_error("shape type " + std::to_string(s.shape) + " is not supported");
    }
    return pref * f;
}

ShapeVar<double> shape_values(const qd_shape &s) { return {s.Ox, s.Oy, s.Oz, s.L1, s.L2, s.L3, s.L4}; }

}  // namespace

cplx shape_transform(const qd_inputs &in, int idx, double Gx0, double Gy0, double Gz0, Cell cell) {
    return shape_transform_t<double>(in, idx, Gx0, Gy0, Gz0, cell, shape_values(in.shapes[idx]));
}

cplx shape_transform_deriv(const qd_inputs &in, int idx, double Gx0, double Gy0, double Gz0, Cell cell, int param) {
    const qd_shape &s = in.shapes[idx];
    ShapeVar<Dual> g{Dual(s.Ox), Dual(s.Oy), Dual(s.Oz), Dual(s.L1), Dual(s.L2), Dual(s.L3), Dual(s.L4)};
    switch (param) {
    case QD_DP_L1: g.L1.d = 1.0; break;
    case QD_DP_L2: g.L2.d = 1.0; break;
    case QD_DP_L3: g.L3.d = 1.0; break;
    case QD_DP_L4: g.L4.d = 1.0; break;
    case QD_DP_L12: g.L1.d = 1.0; g.L2.d = 1.0; break;
    case QD_DP_OX: g.Ox.d = 1.0; break;
    case QD_DP_OY: g.Oy.d = 1.0; break;
    case QD_DP_OZ: g.Oz.d = 1.0; break;
    default: throw std::runtime_error("not a shape-geometry parameter");
    }
    // the quantum well's and wire's direction is selected by Length 1: it is not a length there
    if ((s.shape == 17 || s.shape == 18) && (param == QD_DP_L1 || param == QD_DP_L12)) g.L1.d = 0.0;
    return shape_transform_t<Dual>(in, idx, Gx0, Gy0, Gz0, cell, g).d;
}

namespace {
void run_threads(long n, unsigned nt, const std::function<void(long, long)> &body);
}

// Items of unknown, typically large cost (shape transforms with quadrature): threads from 2048 on.
void parallel_for(long n, const std::function<void(long, long)> &body) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    if (nt > 16) nt = 16;
    if (n < 2048 || nt == 1) { body(0, n); return; }
    run_threads(n, nt, body);
}

// Threads only where each gets at least ~64k units of work (a thread start costs ~20 us).
void parallel_for_cost(long n, double cost, const std::function<void(long, long)> &body) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    if (nt > 16) nt = 16;
    const double want = (double)n * std::max(cost, 1.0) / 65536.0;
    if (want < (double)nt) nt = (unsigned)std::max(1.0, std::floor(want));
    if ((long)nt > n) nt = (unsigned)std::max(1L, n);
    if (nt <= 1) { body(0, n); return; }
    run_threads(n, nt, body);
}

namespace {
void run_threads(long n, unsigned nt, const std::function<void(long, long)> &body) {
    std::vector<std::thread> th;
    std::exception_ptr err;
    std::mutex mu;
    const long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const long a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&, a, b] {
            try { body(a, b); } catch (...) { std::lock_guard<std::mutex> g(mu); err = std::current_exception(); }
        });
    }
    for (auto &t : th) t.join();
    if (err) std::rethrow_exception(err);
}
}  // namespace

void shape_grid_deriv(const qd_inputs &in, int idx, const std::vector<double> &gx, const std::vector<double> &gy,
                      const std::vector<double> &gz, Cell cell, int param, std::vector<cplx> &out) {
    const long ny = gy.size(), nz = gz.size(), n = (long)gx.size() * ny * nz;
    out.assign(n, cplx(0, 0));
    parallel_for(n, [&](long a, long b) {
        for (long i = a; i < b; ++i) {
            const long ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
            out[i] = shape_transform_deriv(in, idx, gx[ix], gy[iy], gz[iz], cell, param);
        }
    });
}

void shape_grid(const qd_inputs &in, int idx, const std::vector<double> &gx, const std::vector<double> &gy,
                const std::vector<double> &gz, Cell cell, std::vector<cplx> &out) {
    const long ny = gy.size(), nz = gz.size(), n = (long)gx.size() * ny * nz;
    out.assign(n, cplx(0, 0));
    parallel_for(n, [&](long a, long b) {
        for (long i = a; i < b; ++i) {
            const long ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
            out[i] = shape_transform(in, idx, gx[ix], gy[iy], gz[iz], cell);
        }
    });
}

}  // namespace qd
