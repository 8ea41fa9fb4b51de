// parameters.in / supercell.in / -key=value readers and the host part of the C ABI.
// File formats follow read_parameters.c:40-502 (flat "key: value" YAML), read_supercell.c:90-303
// (per-shape blocks; lengths nm -> Angstrom; relative strain / Coulomb lengths) and
// read_command_line_input.c:34-59 (-key=value overrides of parameters.in keys).  This reader
// needs no libyaml: the files are flat scalar mappings.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

#include "qd_host.hpp"

namespace qd {

void build_tables(const qd_inputs &in, qd_tables &t);

namespace {
thread_local std::string g_err;

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Ordered "key: value" pairs with '#' comments removed.
std::vector<std::pair<std::string, std::string>> read_pairs(const char *path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error(std::string("failed to open ") + path);
    std::vector<std::pair<std::string, std::string>> out;
    std::string line;
    while (std::getline(f, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        const size_t c = line.find(':');
        if (c == std::string::npos) continue;
        std::string k = trim(line.substr(0, c)), v = trim(line.substr(c + 1));
        if (k.empty()) continue;
        out.emplace_back(k, v);
    }
    return out;
}

double num(const std::string &k, const std::string &v) {
    char *end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (end == v.c_str()) throw std::runtime_error("non-numeric value for key '" + k + "': " + v);
    return d;
}

void apply_parameter(qd_inputs &in, const std::string &k, const std::string &v) {
    const double x = num(k, v);
    const int ix = (int)std::lround(x);
    if (k == "N_shape") in.N_shape = ix;
    else if (k == "M_Gx") in.N[0] = 2 * ix + 1;
    else if (k == "M_Gy") in.N[1] = 2 * ix + 1;
    else if (k == "M_Gz") in.N[2] = 2 * ix + 1;
    else if (k == "Strain_supercell_type") in.strain_type = ix;
    else if (k == "M_S_Gx") in.N_S[0] = 2 * ix + 1;
    else if (k == "M_S_Gy") in.N_S[1] = 2 * ix + 1;
    else if (k == "M_S_Gz") in.N_S[2] = 2 * ix + 1;
    else if (k == "N_VB") in.N_vb = ix;
    else if (k == "N_CB") in.N_cb = ix;
    else if (k == "T") in.T = x;
    else if (k == "Fx") in.F[0] = 1.0e-5 * x;
    else if (k == "Fy") in.F[1] = 1.0e-5 * x;
    else if (k == "Fz") in.F[2] = 1.0e-5 * x;
    else if (k == "Kx") in.K[0] = 0.1 * x;
    else if (k == "Ky") in.K[1] = 0.1 * x;
    else if (k == "Kz") in.K[2] = 0.1 * x;
    else if (k == "Strain") in.strain = ix;
    else if (k == "Isotropic") in.isotropic = ix;
    else if (k == "Piezo") in.piezo = ix;
    else if (k == "CME") in.cme = ix;
    else if (k == "MME") in.mme = ix;
    else if (k == "Overlaps") in.overlaps = ix;
    else if (k == "Localisation") in.localisation = ix;
    else if (k == "Coulomb_supercell_type") in.coulomb_type = ix;
    else if (k == "M_C_Gx") in.N_C[0] = 2 * ix + 1;
    else if (k == "M_C_Gy") in.N_C[1] = 2 * ix + 1;
    else if (k == "M_C_Gz") in.N_C[2] = 2 * ix + 1;
    else if (k == "Softening") in.softening = ix;
    else if (k == "Sigma_softening") in.sigma_softening = 10.0 * x;
    else if (k == "Madelung_ranks") in.madelung_ranks = ix;
    else if (k == "N_line") in.N_line = ix;
    else if (k == "Cross_sections") in.cross_sections = ix;
    else if (k == "N_cross") in.N_cross = ix;
    else if (k == "3D_output") in.out3d = ix;
    else if (k == "N_3D") in.N_3D = ix;
    else if (k.rfind("Linescan_", 0) == 0 || k.rfind("Cross_section_", 0) == 0) {
        static const char *const kLine[7] = {"100", "010", "001", "110", "101", "011", "111"};
        static const char *const kCross[8] = {"100", "010", "001", "m110", "110", "m101", "0m11", "111"};
        const bool line = k[0] == 'L';
        const std::string d = k.substr(line ? 9 : 14);
        for (int i = 0; i < (line ? 7 : 8); ++i)
            if (d == (line ? kLine[i] : kCross[i])) (line ? in.linescan[i] : in.cross_section[i]) = ix;
    }
    else if (k == "Elastic_const") {
        if (ix != 0) throw std::runtime_error("Elastic_const: 1 (position-dependent elastic constants) is not supported");
    }
    // other keys (output toggles, eigensolver workspace) do not affect the solve
}

}  // namespace

void set_error(const std::string &msg) { g_err = msg; }

}  // namespace qd

using namespace qd;

extern "C" const char *qd_last_error(void) { return g_err.c_str(); }

extern "C" int qd_read_inputs(const char *parameters_path, const char *supercell_path, int argc,
                              const char *const *argv, qd_inputs *out) {
    if (!parameters_path || !supercell_path || !out) { set_error("null argument"); return QD_E_ARG; }
    try {
        qd_inputs in;
        std::memset(&in, 0, sizeof(in));
        in.N_shape = 1;
        in.N[0] = in.N[1] = in.N[2] = 1;
        in.N_S[0] = in.N_S[1] = in.N_S[2] = 1;
        in.N_C[0] = in.N_C[1] = in.N_C[2] = 1;
        in.T = 300.0;
        in.sigma_softening = 1.0;
        in.madelung_ranks = 4;
        for (auto &kv : read_pairs(parameters_path)) apply_parameter(in, kv.first, kv.second);
        for (int i = 0; i < argc; ++i) {
            std::string a = argv[i] ? argv[i] : "";
            if (a.empty() || a[0] != '-') throw std::runtime_error("Invalid argument format: " + a + " (must begin with '-')");
            const size_t e = a.find('=');
            if (e == std::string::npos) throw std::runtime_error("Invalid argument (missing '='): " + a);
            apply_parameter(in, a.substr(1, e - 1), a.substr(e + 1));
        }
        if (in.N_shape < 1) throw std::runtime_error("N_shape must be >= 1");
        auto pairs = read_pairs(supercell_path);
        std::vector<qd_shape> sh;
        sh.emplace_back();
        std::memset(&sh.back(), 0, sizeof(qd_shape));
        int RL[3] = {1, 1, 1}, RCL[3] = {1, 1, 1};
        double SL[3] = {0, 0, 0}, CL[3] = {0, 0, 0};
        for (auto &kv : pairs) {
            const std::string &k = kv.first;
            const double v = num(k, kv.second);
            if ((int)sh.size() > in.N_shape)
                throw std::runtime_error("N_shape is less than the number of shapes given in supercell.in");
            qd_shape &c = sh.back();
            const bool cell = sh.size() == 1;
            if (k == "Embedded") c.embedded = (int)std::lround(v);
            else if (k == "QD shape") c.shape = (int)std::lround(v);
            else if (k == "Material") c.material = (int)std::lround(v);
            else if (k == "Compos x") c.x = v;
            else if (k == "Compos y") c.y = v;
            else if (k == "Centre x") c.Ox = 10 * v;
            else if (k == "Centre y") c.Oy = 10 * v;
            else if (k == "Centre z") c.Oz = 10 * v;
            else if (k == "Length 1") c.L1 = 10 * v;
            else if (k == "Length 2") c.L2 = 10 * v;
            else if (k == "Length 3") c.L3 = 10 * v;
            else if (k == "Length 4") c.L4 = 10 * v;
            else if (k == "Length 5") c.L5 = 10 * v;
            else if (k.rfind("Strain length ", 0) == 0 || k.rfind("Coulomb length ", 0) == 0) {
                const int a = k.back() - '1';
                if (a < 0 || a > 2) throw std::runtime_error("bad key " + k);
                const bool strain = k[0] == 'S';
                const double Lc = a == 0 ? c.L1 : (a == 1 ? c.L2 : c.L3);
                const int type = strain ? in.strain_type : in.coulomb_type;
                double &dst = strain ? SL[a] : CL[a];
                int &rel = strain ? RL[a] : RCL[a];
                if (cell) {
                    if (type == 0) dst = 10 * v;
                    else { rel = (int)std::lround(v); dst = Lc * rel; }
                }
                if (!strain && a == 2 && cell) { sh.emplace_back(); std::memset(&sh.back(), 0, sizeof(qd_shape)); }
            } else if (k == "Euler alpha angle") c.alpha = v;
            else if (k == "Euler beta angle") c.beta = v;
            else if (k == "Euler gamma angle") {
                c.gamma = v;
                sh.emplace_back();
                std::memset(&sh.back(), 0, sizeof(qd_shape));
            }
        }
        // the trailing (empty) entry opened after the last complete block
        if ((int)sh.size() == in.N_shape + 1) sh.pop_back();
        if ((int)sh.size() != in.N_shape)
            throw std::runtime_error("N_shape (" + std::to_string(in.N_shape) + ") does not match the " +
                                     std::to_string(sh.size()) + " shapes in supercell.in");
        for (int a = 0; a < 3; ++a) { in.RL[a] = RL[a]; in.RCL[a] = RCL[a]; in.SL[a] = SL[a]; in.CL[a] = CL[a]; }
        if (in.strain && (in.SL[0] <= 0 || in.SL[1] <= 0 || in.SL[2] <= 0))
            throw std::runtime_error("strain supercell lengths missing");
        in.shapes = (qd_shape *)std::malloc(sizeof(qd_shape) * in.N_shape);
        std::memcpy(in.shapes, sh.data(), sizeof(qd_shape) * in.N_shape);
        *out = in;
        return QD_OK;
    } catch (const std::exception &e) {
        set_error(e.what());
        return QD_E_INPUT;
    }
}

extern "C" void qd_free_inputs(qd_inputs *in) {
    if (in && in->shapes) { std::free(in->shapes); in->shapes = nullptr; }
}

extern "C" int qd_window(const qd_inputs *in, int *il, int *iu) {
    if (!in || !il || !iu) { set_error("null argument"); return QD_E_ARG; }
    const long NG = (long)in->N[0] * in->N[1] * in->N[2];
    const long cb_min = 6 * NG;
    *il = (int)(cb_min - 2 * in->N_vb + 1);
    *iu = (int)(cb_min + 2 * in->N_cb);
    if (*il < 1 || *iu > 8 * NG) { set_error("eigen window outside the spectrum"); return QD_E_ARG; }
    return QD_OK;
}

namespace {
// The builder's intermediates kept with the tables, and the inputs they were built from.
struct RetainedState {
    TableState st;
    qd_inputs in;
    std::vector<qd_shape> shapes;
};

bool same_inputs(const qd_inputs &a, const qd_inputs &b, const std::vector<qd_shape> &bs) {
    if (a.N_shape != b.N_shape || (int)bs.size() != a.N_shape || !a.shapes) return false;
    for (int k = 0; k < 3; ++k)
        if (a.N[k] != b.N[k] || a.N_S[k] != b.N_S[k] || a.RL[k] != b.RL[k] || a.SL[k] != b.SL[k] || a.F[k] != b.F[k] ||
            a.K[k] != b.K[k])
            return false;
    if (a.strain_type != b.strain_type || a.T != b.T || a.strain != b.strain || a.isotropic != b.isotropic ||
        a.piezo != b.piezo || a.softening != b.softening || a.sigma_softening != b.sigma_softening || a.N_vb != b.N_vb ||
        a.N_cb != b.N_cb)
        return false;
    for (int s = 0; s < a.N_shape; ++s) {
        const qd_shape &x = a.shapes[s], &y = bs[s];
        if (x.shape != y.shape || x.embedded != y.embedded || x.material != y.material || x.x != y.x || x.y != y.y ||
            x.Ox != y.Ox || x.Oy != y.Oy || x.Oz != y.Oz || x.L1 != y.L1 || x.L2 != y.L2 || x.L3 != y.L3 || x.L4 != y.L4 ||
            x.L5 != y.L5 || x.alpha != y.alpha || x.beta != y.beta || x.gamma != y.gamma)
            return false;
    }
    return true;
}

int tables_build_impl(const qd_inputs *in, qd_tables *out, bool retain) {
    if (!in || !out || !in->shapes || in->N_shape < 1) { set_error("null or empty inputs"); return QD_E_ARG; }
    for (int a = 0; a < 3; ++a)
        if (in->N[a] < 1 || (in->N[a] % 2) == 0) { set_error("plane-wave counts must be odd and positive"); return QD_E_ARG; }
    std::memset(out, 0, sizeof(*out));
    try {
        if (retain) {
            std::unique_ptr<RetainedState> r(new RetainedState());
            build_state(*in, r->st);
            compose_tables(*in, r->st, *out);
            r->in = *in;
            r->shapes.assign(in->shapes, in->shapes + in->N_shape);
            r->in.shapes = nullptr;
            out->state = r.release();
        } else {
            build_tables(*in, *out);
        }
    } catch (const std::exception &e) {
        qd_tables_free(out);
        set_error(e.what());
        const std::string m = e.what();
        return m.find("not supported") != std::string::npos ? QD_E_UNSUPPORTED : QD_E_INPUT;
    }
    const size_t n = (size_t)QD_NTAB * out->NT[0] * out->NT[1] * out->NT[2] * 2;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(out->tab[i])) { qd_tables_free(out); set_error("non-finite value in coefficient tables"); return QD_E_NAN; }
    return QD_OK;
}
}  // namespace

extern "C" int qd_tables_build(const qd_inputs *in, qd_tables *out) { return tables_build_impl(in, out, false); }

extern "C" int qd_tables_build_retain(const qd_inputs *in, qd_tables *out) { return tables_build_impl(in, out, true); }

extern "C" void qd_tables_free(qd_tables *t) {
    if (!t) return;
    for (int a = 0; a < 3; ++a) { std::free(t->G[a]); t->G[a] = nullptr; }
    std::free(t->tab);
    t->tab = nullptr;
    delete static_cast<RetainedState *>(t->state);
    t->state = nullptr;
}

// dL/dtheta for L = Re sum_t sum_d W[t][d] T_t[d]: one reverse pass through the table builder
// (TableState / tables_vjp) gives dL/dchi_s and dL/d(material fields); geometry parameters then
// contract dL/dchi_s with the exact d chi_s / d theta (dual numbers through the shape transforms),
// compositions contract dL/d(fields) with d(fields)/dx, the field is linear.
namespace qd {

int vjp_check_params(const qd_inputs *in, const double *W, int nparams, const int *params, const unsigned *masks,
                     const double *grad) {
    if (!in || !W || !in->shapes || nparams < 1 || !params || !masks || !grad) { set_error("bad argument"); return QD_E_ARG; }
    for (int j = 0; j < nparams; ++j) {
        if (params[j] < QD_DP_FX || params[j] > QD_DP_L4) { set_error("unknown parameter"); return QD_E_ARG; }
        if (params[j] > QD_DP_FZ) {
            if (masks[j] & 1u) { set_error("shape_mask bit 0 (the matrix) is not a differentiable shape"); return QD_E_ARG; }
            bool any = false;
            for (int s = 1; s < in->N_shape && s < 32; ++s) {
                if (!(masks[j] & (1u << s))) continue;
                any = true;
                // the pyramid's faces are one face segment turned by 2 atan(L2 / L1) k: L1 or L2
                // alone turns them off the axes where the transform takes its limiting branches,
                // which carry no derivative across the axis
                if ((in->shapes[s].shape == 5 || in->shapes[s].shape == 6) &&
                    (params[j] == QD_DP_L1 || params[j] == QD_DP_L2)) {
                    set_error("Length 1 or Length 2 alone of a rectangular-base pyramid (shapes 5, 6) is not supported; "
                              "use QD_DP_L12");
                    return QD_E_UNSUPPORTED;
                }
            }
            if (!any) { set_error("shape_mask selects no shape (bit s = shape s >= 1)"); return QD_E_ARG; }
        }
    }
    return QD_OK;
}

void vjp_contract(const qd_inputs &in, const TableState &st, const TableGrad &g, int nparams, const int *params,
                  const unsigned *masks, double *grad) {
    for (int j = 0; j < nparams; ++j) {
        const int p = params[j];
        double acc = 0.0;
        if (p <= QD_DP_FZ) {
            acc = g.gF[p - QD_DP_FX];
        } else if (p == QD_DP_X) {
            for (int s = 1; s < in.N_shape && s < 32; ++s) {
                if (!(masks[j] & (1u << s))) continue;
                const Material d = material_dx(in, s);
                const Material &gm = g.gmp[s];
                acc += gm.a * d.a + gm.e14 * d.e14 + gm.B114 * d.B114 + gm.B124 * d.B124 + gm.B156 * d.B156 +
                       gm.nr * d.nr + gm.sc * d.sc + gm.g1 * d.g1 + gm.g2 * d.g2 + gm.g3 * d.g3 + gm.P * d.P +
                       gm.ac * d.ac + gm.av * d.av + gm.b * d.b + gm.d * d.d + gm.E_cb * d.E_cb +
                       gm.E_hh * d.E_hh + gm.E_lh * d.E_lh + gm.E_so * d.E_so;
            }
        } else if (j < (int)g.gparam.size() && std::isfinite(g.gparam[j])) {
            acc = g.gparam[j];
        } else if ((p == QD_DP_OX || p == QD_DP_OY || p == QD_DP_OZ) && !g.gcentre.empty()) {
            for (int s = 1; s < in.N_shape && s < 32; ++s)
                if (masks[j] & (1u << s)) acc += g.gcentre[3 * s + (p - QD_DP_OX)];
        } else {
            for (int s = 1; s < in.N_shape && s < 32; ++s) {
                if (!(masks[j] & (1u << s))) continue;
                std::vector<cplx> dchi;
                shape_grid_deriv(in, s, st.gE[0], st.gE[1], st.gE[2], Cell::Electronic, p, dchi);
                for (size_t i = 0; i < dchi.size(); ++i)
                    acc += g.gchiE[s].v[i].real() * dchi[i].real() - g.gchiE[s].v[i].imag() * dchi[i].imag();
                if (st.use_strain && !st.same) {
                    shape_grid_deriv(in, s, st.gS[0], st.gS[1], st.gS[2], Cell::Strain, p, dchi);
                    for (size_t i = 0; i < dchi.size(); ++i)
                        acc += g.gchiS[s].v[i].real() * dchi[i].real() - g.gchiS[s].v[i].imag() * dchi[i].imag();
                }
            }
        }
        grad[j] = acc;
    }
}

}  // namespace qd

namespace {
int tables_vjp_impl(const qd_tables *t, const qd_inputs *in, const double *W, int nparams, const int *params,
                    const unsigned *masks, double *grad);
}

extern "C" int qd_tables_vjp(const qd_inputs *in, const double *W, int nparams, const int *params,
                             const unsigned *masks, double *grad) {
    return tables_vjp_impl(nullptr, in, W, nparams, params, masks, grad);
}

extern "C" int qd_tables_vjp_retained(const qd_tables *t, const qd_inputs *in, const double *W, int nparams,
                                      const int *params, const unsigned *masks, double *grad) {
    return tables_vjp_impl(t, in, W, nparams, params, masks, grad);
}

namespace {
int tables_vjp_impl(const qd_tables *t, const qd_inputs *in, const double *W, int nparams, const int *params,
                    const unsigned *masks, double *grad) {
    const int rc = vjp_check_params(in, W, nparams, params, masks, grad);
    if (rc != QD_OK) return rc;
    try {
        // the intermediates kept with the tables when they were built from these same inputs,
        // else a fresh forward pass
        const RetainedState *r = t ? static_cast<const RetainedState *>(t->state) : nullptr;
        TableState fresh;
        const TableState *stp = &fresh;
        if (r && same_inputs(*in, r->in, r->shapes)) stp = &r->st;
        else build_state(*in, fresh);
        const TableState &st = *stp;
        TableGrad g;
        tables_vjp(*in, st, reinterpret_cast<const cplx *>(W), g);
        vjp_contract(*in, st, g, nparams, params, masks, grad);
    } catch (const std::exception &e) {
        set_error(e.what());
        const std::string m = e.what();
        return m.find("not supported") != std::string::npos ? QD_E_UNSUPPORTED : QD_E_INPUT;
    }
    return QD_OK;
}

}  // namespace

// Test support: shape transforms chi_idx(G) of one shape at n wave vectors (host), and their
// derivatives with respect to one geometry parameter (param >= QD_DP_L1) when param is not -1.
extern "C" int qd_selftest_shape_host(const qd_inputs *in, int idx, int cell, int param, int n, const double *G,
                                      double *out) {
    if (!in || !in->shapes || !G || !out || n < 0 || idx < 0 || idx >= in->N_shape) { set_error("bad argument"); return QD_E_ARG; }
    if (cell != 0 && cell != 1 && cell != 3) { set_error("cell must be 0, 1 or 3"); return QD_E_ARG; }
    try {
        const Cell c = cell == 0 ? Cell::Electronic : (cell == 1 ? Cell::Strain : Cell::Coulomb);
        for (int i = 0; i < n; ++i) {
            const cplx v = param < 0 ? shape_transform(*in, idx, G[3 * i], G[3 * i + 1], G[3 * i + 2], c)
                                     : shape_transform_deriv(*in, idx, G[3 * i], G[3 * i + 1], G[3 * i + 2], c, param);
            out[2 * i] = v.real();
            out[2 * i + 1] = v.imag();
        }
    } catch (const std::exception &e) {
        set_error(e.what());
        const std::string m = e.what();
        return m.find("not supported") != std::string::npos ? QD_E_UNSUPPORTED : QD_E_INPUT;
    }
    return QD_OK;
}

// dT/dtheta of the Delta-G tables.  The field enters the four band-edge tables through the
// transform of the sawtooth potential (linear in F); shape lengths and compositions enter
// through the shape transforms, strain and piezo fields, so they are differentiated by central
// differences of the whole FP64 builder.
extern "C" int qd_tables_derivative(const qd_inputs *in, int param, unsigned shape_mask, double rel_step, double *dT) {
    if (!in || !dT || !in->shapes || param < QD_DP_FX || param > QD_DP_L4) { set_error("bad argument"); return QD_E_ARG; }
    int NT[3];
    for (int a = 0; a < 3; ++a) NT[a] = 2 * in->N[a] - 1;
    const size_t nT = (size_t)NT[0] * NT[1] * NT[2];
    std::memset(dT, 0, sizeof(double) * 2 * QD_NTAB * nT);
    if (param <= QD_DP_FZ) {
        const int a = param - QD_DP_FX;
        const double La = a == 0 ? in->shapes[0].L1 : (a == 1 ? in->shapes[0].L2 : in->shapes[0].L3);
        const int c[3] = {NT[0] / 2, NT[1] / 2, NT[2] / 2};
        for (int m = 0; m < NT[a]; ++m) {
            const int dd = m - c[a];
            if (dd == 0) continue;
            int idx[3] = {c[0], c[1], c[2]};
            idx[a] = m;
            const size_t i = ((size_t)idx[0] * NT[1] + idx[1]) * NT[2] + idx[2];
            const double sign = (dd % 2 == 0) ? -1.0 : 1.0;
            // i sign / (-dd 2 pi / L): purely imaginary
            const double im = sign / (-(double)dd * 2.0 * 3.14159265358979323846 / La);
            for (int t = QD_T_ECB; t <= QD_T_ESO; ++t) dT[2 * ((size_t)t * nT + i) + 1] = im;
        }
        return QD_OK;
    }
    if (rel_step <= 0) { set_error("rel_step must be positive"); return QD_E_ARG; }
    int first = -1;
    for (int s = 1; s < in->N_shape && s < 32; ++s)
        if (shape_mask & (1u << s)) { first = s; break; }
    if (first < 0) { set_error("shape_mask selects no shape (bit s = shape s >= 1)"); return QD_E_ARG; }
    auto value = [&](const qd_shape &sh) {
        switch (param) {
        case QD_DP_L1: case QD_DP_L12: return sh.L1;
        case QD_DP_L2: return sh.L2;
        case QD_DP_L3: return sh.L3;
        case QD_DP_L4: return sh.L4;
        default: return sh.x;
        }
    };
    // compositions and positions step by rel_step itself (x in [0, 1]; positions in Angstrom)
    const bool absolute = param == QD_DP_X || (param >= QD_DP_OX && param <= QD_DP_OZ);
    const double h = absolute ? rel_step : rel_step * std::fabs(value(in->shapes[first]));
    if (h <= 0) { set_error("zero step (parameter value is zero)"); return QD_E_ARG; }
    std::vector<qd_shape> sp(in->shapes, in->shapes + in->N_shape), sm = sp;
    for (int s = 1; s < in->N_shape && s < 32; ++s) {
        if (!(shape_mask & (1u << s))) continue;
        for (int sgn = 0; sgn < 2; ++sgn) {
            qd_shape &q = sgn == 0 ? sp[s] : sm[s];
            const double d = sgn == 0 ? h : -h;
            switch (param) {
            case QD_DP_L1: q.L1 += d; break;
            case QD_DP_L2: q.L2 += d; break;
            case QD_DP_L3: q.L3 += d; break;
            case QD_DP_L4: q.L4 += d; break;
            case QD_DP_L12: q.L1 += d; q.L2 += d; break;
            case QD_DP_OX: q.Ox += d; break;
            case QD_DP_OY: q.Oy += d; break;
            case QD_DP_OZ: q.Oz += d; break;
            default: q.x += d; break;
            }
        }
    }
    qd_inputs ip = *in, im = *in;
    ip.shapes = sp.data();
    im.shapes = sm.data();
    qd_tables tp{}, tm{};
    int rc = qd_tables_build(&ip, &tp);
    if (rc != QD_OK) return rc;
    rc = qd_tables_build(&im, &tm);
    if (rc != QD_OK) { qd_tables_free(&tp); return rc; }
    const double inv = 1.0 / (2.0 * h);
    for (size_t i = 0; i < 2 * QD_NTAB * nT; ++i) dT[i] = (tp.tab[i] - tm.tab[i]) * inv;
    qd_tables_free(&tp);
    qd_tables_free(&tm);
    return QD_OK;
}
