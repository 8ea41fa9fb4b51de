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

template <class F>
cplx gl_apply(const GLRule &r, const F &f, double a, double b) {
    const double h = 0.5 * (b - a), c = 0.5 * (b + a);
    cplx s = 0;
    for (size_t i = 0; i < r.x.size(); ++i) s += r.w[i] * f(c + h * r.x[i]);
    return h * s;
}

template <class F>
cplx integrate(const F &f, double a, double b, int depth = 0) {
    const cplx lo = gl_apply(gl_lo(), f, a, b);
    const cplx hi = gl_apply(gl_hi(), f, a, b);
    if (std::abs(hi - lo) <= 1e-14 * std::max(1.0, std::abs(hi)) || depth >= 12) return hi;
    const double m = 0.5 * (a + b);
    return integrate(f, a, m, depth + 1) + integrate(f, m, b, depth + 1);
}

// int_0^1 x cos(alpha sqrt(1-x^2)) J0(beta x) dx
double int_J0(double alpha, double beta) {
    auto f = [&](double x) {
        const double u = 1.0 - x * x;
        return cplx(x * std::cos(alpha * std::sqrt(u > 0 ? u : 0.0)) * ::j0(beta * x), 0.0);
    };
    return integrate(f, 0.0, 1.0).real();
}

// int_{R1/(R2-R1)}^{R2/(R2-R1)} x J0((R2-R1) alpha x) e^{i beta x} dx
cplx int_TC(double alpha, double R1, double R2, double beta) {
    auto f = [&](double x) { return x * ::j0((R2 - R1) * alpha * x) * std::exp(cplx(0.0, beta * x)); };
    return integrate(f, R1 / (R2 - R1), R2 / (R2 - R1));
}

// int over the same limits of x^2 J0((R2-R1) alpha x)
double int_TC_z(double alpha, double R1, double R2) {
    auto f = [&](double x) { return cplx(x * x * ::j0((R2 - R1) * alpha * x), 0.0); };
    return integrate(f, R1 / (R2 - R1), R2 / (R2 - R1)).real();
}

inline cplx cis_m(double a) { return cplx(std::cos(a), -std::sin(a)); }   // e^{-ia}
inline double sinc_len(double G, double L) {   // L sin(GL/2)/(GL/2)
    const double t = 0.5 * G * L;
    return t == 0.0 ? L : L * std::sin(t) / t;
}
const cplx I1(0.0, 1.0);

// (e^{iGa} - 1)/(iG), a at G = 0 (characteristic_function_fourier_integrals.c:28-50)
cplx edge_integral(double a, double G) {
    if (std::fabs(G) < kEps) return cplx(a, 0);
    return (std::exp(cplx(0, G * a)) - 1.0) / (I1 * G);
}

// Triangular pyramid segment of a rectangular-base pyramid (apex height h over the base plane,
// base edge at x0 with half-width y0), side k in 0..3 (characteristic_function_fourier.c:1116-1271
// with N_side == 0).
cplx pyramid_segment(double Gx0, double Gy0, double Gz, int k, double L1, double L2, double h) {
    double alpha, x0, y0;
    if (k == 0 || k == 2) { alpha = std::atan(L2 / L1); x0 = L1; y0 = L2; }
    else { alpha = std::atan(L1 / L2); x0 = L2; y0 = L1; }
    const double ang = 2.0 * alpha * k;
    const double Gx = std::cos(ang) * Gx0 + std::sin(ang) * Gy0;
    const double Gy = -std::sin(ang) * Gx0 + std::cos(ang) * Gy0;
    const double Gm = std::sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const bool zx = std::fabs(Gx) < kEps, zy = std::fabs(Gy) < kEps, zz = std::fabs(Gz) < kEps;
    if (Gm < kEps) return x0 * y0 * h / 3.0;
    if (zx && zy)
        return (I1 * 2.0 * y0 * x0 / Gz) * ((1.0 - I1 * Gz * h - cis_m(Gz * h)) / (Gz * Gz * h * h) - 0.5);
    if (zx && zz) return (2.0 * h * x0 / (Gy * Gy * y0)) * (1.0 - std::sin(Gy * y0) / (Gy * y0));
    if (zy && zz)
        return (2.0 * h * y0 / (Gx * Gx * x0)) *
               (-1.0 - cis_m(Gx * x0) - 2.0 * I1 / (Gx * x0) + 2.0 * I1 * cis_m(Gx * x0) / (Gx * x0));
    if (zx) {
        const double sp = Gy * y0 + Gz * h, sm = Gy * y0 - Gz * h;
        const cplx pre = (-I1 * x0) / (Gz * Gy);
        const cplx tail = -2.0 * (std::cos(Gy * y0) - 1.0) / (Gy * y0);
        cplx body;
        if (std::fabs(sm) < kEps) body = -I1 + (std::exp(I1 * sp) - 1.0) / sp;
        else if (std::fabs(sp) < kEps) body = (cis_m(sm) - 1.0) / sm + I1;
        else body = (cis_m(sm) - 1.0) / sm + (std::exp(I1 * sp) - 1.0) / sp;
        return pre * (cis_m(Gz * h) * body + tail);
    }
    if (zy) {
        const double q = Gx - Gz * h / x0;
        const cplx base = (cis_m(Gx * x0) * (1.0 + I1 * Gx * x0) - 1.0) / (Gx * Gx);
        const cplx pre = 2.0 * I1 * y0 / (Gz * x0);
        if (std::fabs(q) < kEps) return pre * (0.5 * cis_m(Gz * h) * x0 * x0 - base);
        const double r = Gx * x0 - Gz * h;
        return pre * (cis_m(Gz * h) * (cis_m(r) * (1.0 + I1 * r) - 1.0) / (q * q) - base);
    }
    if (zz) {
        const double pp = Gx + Gy * y0 / x0, pm = Gx - Gy * y0 / x0;
        const cplx Ap = (I1 / (x0 * pp * pp)) * (cis_m(x0 * pp) - 1.0);
        const cplx Am = (I1 / (x0 * pm * pm)) * (cis_m(x0 * pm) - 1.0);
        if (std::fabs(pm) < kEps) return (h / Gy) * (1.0 / pp - Ap - 0.5 * I1 * x0);
        if (std::fabs(pp) < kEps) return (h / Gy) * (-1.0 / pm + Am + 0.5 * I1 * x0);
        return (h / Gy) * (1.0 / pp - 1.0 / pm - Ap + Am);
    }
    const double s1 = Gy * y0 / x0, s2 = Gz * h / x0;
    return (-1.0 / (Gy * Gz)) * (edge_integral(x0, -Gx - s1 + s2) * cis_m(Gz * h) -
                                 edge_integral(x0, -Gx + s1 + s2) * cis_m(Gz * h) -
                                 edge_integral(x0, -Gx - s1) + edge_integral(x0, -Gx + s1));
}

// One of N wedge segments of a regular N-sided prism of inradius x0 and height h centred in z
// (characteristic_function_fourier.c:1276-1396).
cplx prism_segment(double Gx0, double Gy0, double Gz, int k, int N, double x0, double h) {
    const double alpha = kPi / N;
    const double ang = 2.0 * kPi * k / N;
    const double Gx = std::cos(ang) * Gx0 + std::sin(ang) * Gy0;
    const double Gy = -std::sin(ang) * Gx0 + std::cos(ang) * Gy0;
    const double Gm = std::sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    const double ta = std::tan(alpha);
    const bool zx = std::fabs(Gx) < kEps, zy = std::fabs(Gy) < kEps, zz = std::fabs(Gz) < kEps;
    if (Gm < kEps) return x0 * x0 * h * ta;
    if (zx && zy) return 2.0 * x0 * x0 * ta * std::sin(0.5 * Gz * h) / Gz;
    if (zx && zz) return (2.0 * h / (Gy * Gy * ta)) * (1.0 - std::cos(Gy * x0 * ta));
    const cplx xterm = I1 * x0 * cis_m(Gx * x0) + (1.0 / Gx) * (cis_m(Gx * x0) - 1.0);
    if (zy && zz) return (2.0 * h * ta / Gx) * xterm;
    if (zx) return (4.0 * std::sin(0.5 * Gz * h) / (Gy * Gy * Gz * ta)) * (1.0 - std::cos(Gy * x0 * ta));
    if (zy) return (4.0 * std::sin(0.5 * Gz * h) * ta / (Gz * Gx)) * xterm;
    const double up = Gy * ta + Gx, um = -Gy * ta + Gx;
    const cplx ep = 1.0 - std::exp(cplx(0, -Gy * x0 * ta - Gx * x0));
    const cplx em = 1.0 - std::exp(cplx(0, Gy * x0 * ta - Gx * x0));
    cplx face;
    if (std::fabs(up) < kEps) face = I1 * x0 - em / um;
    else if (std::fabs(um) < kEps) face = ep / up - I1 * x0;
    else face = ep / up - em / um;
    if (zz) return (h / Gy) * face;
    return (2.0 / (Gy * Gz)) * std::sin(0.5 * Gz * h) * face;
}

// 4 pi (sin g - g cos g)/g^3 scaled: the unit-sphere transform.
inline double sphere_kernel(double g) { return (std::sin(g) - g * std::cos(g)) / (g * g * g); }

}  // namespace

cplx shape_transform(const qd_inputs &in, int idx, double Gx0, double Gy0, double Gz0, Cell cell) {
    const qd_shape &s = in.shapes[idx];
    const qd_shape &c0 = in.shapes[0];
    double vol;
    if (cell == Cell::Coulomb) vol = in.CL[0] * in.CL[1] * in.CL[2];
    else if (cell == Cell::Electronic || idx == 0) vol = c0.L1 * c0.L2 * c0.L3;
    else vol = in.SL[0] * in.SL[1] * in.SL[2];
    cplx pref = cis_m(Gx0 * s.Ox + Gy0 * s.Oy + Gz0 * s.Oz) / vol;

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
        pref *= std::exp(-0.5 * in.sigma_softening * in.sigma_softening * (Gx * Gx + Gy * Gy + Gz * Gz));

    const double L1 = s.L1, L2 = s.L2, L3 = s.L3, L4 = s.L4;
    cplx f;
    switch (s.shape) {
    case 0:   // cuboid centred on O
        f = sinc_len(Gx, L1) * sinc_len(Gy, L2) * sinc_len(Gz, L3);
        break;
    case 1:   // sphere of radius L1
        f = std::fabs(Gm) < kEps ? 4.0 * kPi * L1 * L1 * L1 / 3.0
                                  : 4.0 * kPi * L1 * L1 * L1 * sphere_kernel(Gm * L1);
        break;
    case 2: { // ellipsoid, semi-axes L1 L2 L3
        const double g = std::sqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2 + Gz * Gz * L3 * L3);
        f = std::fabs(g) < kEps ? 4.0 * kPi * L1 * L2 * L3 / 3.0 : 4.0 * kPi * L1 * L2 * L3 * sphere_kernel(g);
        break;
    }
    case 3: { // hemisphere, base centred on O, radius L1
        const double R = L1;
        const double sph = 2.0 * kPi * R * R * R * sphere_kernel(Gm * R);
        if (std::fabs(Gm) < kEps) f = 2.0 * kPi * R * R * R / 3.0;
        else if (std::fabs(Gp) < kEps && std::fabs(Gz) > kEps) {
            const double a = R * Gz;
            f = sph - (2.0 * kPi * I1 * R * R / Gz) * (0.5 - (a * std::sin(a) + std::cos(a) - 1.0) / (a * a));
        } else if (std::fabs(Gz) < kEps && std::fabs(Gp) > kEps) f = sph;
        else
            f = sph - (2.0 * kPi * I1 * R * ::j1(Gp * R)) / (Gp * Gz) +
                (2.0 * kPi * I1 * R * R * int_J0(Gz * R, Gp * R)) / Gz;
        break;
    }
    case 4: { // lens (half ellipsoid), base centred on O, semi-axes L1 L2 L3
        const double gm = std::sqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2 + Gz * Gz * L3 * L3);
        const double gp = std::sqrt(Gx * Gx * L1 * L1 + Gy * Gy * L2 * L2);
        const double V2 = 2.0 * kPi * L1 * L2 * L3;
        if (std::fabs(gm) < kEps) { f = V2 / 3.0; break; }
        const double sph = V2 * sphere_kernel(gm);
        const double a = Gz * L3;
        if (std::fabs(gp) < kEps)
            f = sph - (2.0 * kPi * I1 * L1 * L2 / Gz) * (0.5 - (std::cos(a) + a * std::sin(a) - 1.0) / (a * a));
        else if (std::fabs(Gz) < kEps) f = sph;
        else f = sph - (2.0 * kPi * I1 * L1 * L2 / Gz) * (::j1(gp) / gp - int_J0(a, gp));
        break;
    }
    case 6: { // truncated rectangular pyramid = full pyramid minus its top
        const double h_full = L4 * L1 / (L1 - L3);
        cplx acc = 0;
        for (int k = 0; k < 4; ++k) acc += pyramid_segment(Gx, Gy, Gz, k, L1, L2, h_full);
        const double t1 = L3, t2 = L2 * L3 / L1, th = L4 * L3 / (L1 - L3);
        cplx top = 0;
        for (int k = 0; k < 4; ++k) top += pyramid_segment(Gx, Gy, Gz, k, t1, t2, th);
        f = acc - cis_m(Gz * L4) * top;
        break;
    }
    case 12: { // truncated cone: base radius L1 at O, top radius L2 at height L3
        const double R1 = L1, R2 = L2, H = L3, dR = L1 - L2;
        if (std::fabs(Gm) < kEps) { f = kPi * H * (R1 * R1 + R1 * R2 + R2 * R2) / 3.0; break; }
        if (std::fabs(Gp) < kEps) {
            const double a = Gz * H;
            f = (2.0 * I1 * kPi / Gz) *
                (0.5 * R2 * R2 * (cis_m(a) - 1.0) +
                 (dR * dR / (a * a)) * (cis_m(a) * (I1 * Gz * R2 * H / dR - 1.0) + (1.0 - I1 * R1 * Gz * H / dR)) -
                 0.5 * (R1 * R1 - R2 * R2));
            break;
        }
        const double jj = (R1 * ::j1(Gp * R1) - R2 * ::j1(Gp * R2)) / Gp;
        if (std::fabs(Gz) < kEps) {
            f = (2.0 * kPi * H) * (R2 * ::j1(Gp * R2) / Gp - dR * dR * int_TC_z(Gp, R2, R1) + (R1 / dR) * jj);
        } else {
            f = (2.0 * I1 * kPi / Gz) * (R2 * (cis_m(Gz * H) - 1.0) * ::j1(Gp * R2) / Gp +
                                         dR * dR * cis_m(Gz * H * R1 / dR) * int_TC(Gp, R2, R1, Gz * H) - jj);
        }
        break;
    }
    case 16: { // regular prism: inradius L1, height L2, L3/10 sides
        const int nside = (int)std::lround(0.1 * L3);
        cplx acc = 0;
        for (int k = 0; k < nside; ++k) acc += prism_segment(Gx, Gy, Gz, k, nside, L1, L2);
        f = acc;
        break;
    }
    case 17: { // quantum well along axis lround(L1/10), thickness L2, spanning the cell otherwise
        const int d = (int)std::lround(0.1 * L1);
        double Ls[3];
        if (cell == Cell::Electronic) { Ls[0] = c0.L1; Ls[1] = c0.L2; Ls[2] = c0.L3; }
        else { Ls[0] = in.SL[0]; Ls[1] = in.SL[1]; Ls[2] = in.SL[2]; }
        if (d < 0 || d > 2) throw std::runtime_error("quantum-well direction must be 0, 1 or 2");
        Ls[d] = L2;
        f = sinc_len(Gx, Ls[0]) * sinc_len(Gy, Ls[1]) * sinc_len(Gz, Ls[2]);
        break;
    }
    default:
        throw std::runtime_error("shape type " + std::to_string(s.shape) + " is not supported");
    }
    return pref * f;
}

void parallel_for(long n, const std::function<void(long, long)> &body) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    if (nt > 16) nt = 16;
    if (n < 2048 || nt == 1) { body(0, n); return; }
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
