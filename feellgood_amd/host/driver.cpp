// driver.cpp -- fgmi_run, the drop-in time loop: reads FeeLLGood settings (YAML or JSON, a file
// or "-" for stdin) and the .msh mesh, runs the reference's integration loop on the gfx950
// engine through the C-ABI, and writes the .evol and .sol outputs.
//
// Follows src/main.cpp:103-232 (options, random seed), src/settings.cpp:314-616 (settings and
// defaults from src/default-settings.yml), src/time_integration.cpp:11-38,136-254 (TimeStepper
// and the loop with its retry semantics) and src/save.cpp:13-186 (.evol columns, .sol files).
// initial_magnetization and Bext take the reference's three expressions (of x, y, z in metres and
// of t in seconds, src/settings.cpp:454-543), or a function definition whose body returns them,
// and Bext as space and time functions (H = time(t) / mu0 * space(x) at the Gauss points), through
// fgh::Expr. Spin accumulation (spin_accumulation.enable) solves the potential once and the spin
// accumulation at every compute_all, with its field in the element pass and its columns in the
// .sol files. Not supported, reported as errors: function bodies beyond one return statement,
// functions of the regions, region-qualified evol columns, Thiele_length and recentering.
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <csignal>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/fgmi.h"
#include "../../include/fgmi_host.h"
#include "expr.h"
#include "settings.h"

namespace {

// SIGINT and SIGTERM set this flag; the time loop saves the state and terminates at the start of
// its next step attempt (src/main.cpp:15-18,256-271; src/time_integration.cpp:99-134)
volatile sig_atomic_t received_signal = 0;
void signal_handler(int sig) { received_signal = sig; }

// The user's defaults file, read before the settings (src/main.cpp:20-34,226):
// $XDG_CONFIG_HOME/feellgood/default.yml, or $HOME/.config/feellgood/default.yml
std::string user_defaults_filename()
{
    const char *x = getenv("XDG_CONFIG_HOME");
    std::string home = (x && *x) ? x : "";
    if (home.empty()) {
        const char *h = getenv("HOME");
        home = std::string(h ? h : "") + "/.config";
    }
    return home + "/feellgood/default.yml";
}

// Settings over defaults: maps merge key by key (the later file's value wins, nested maps merged
// the same way), anything else is replaced -- the effect of reading the two files one after the
// other into the same Settings (src/settings.cpp:596-616).
fgh::Node merge_settings(fgh::Node base, const fgh::Node &over)
{
    if (base.kind != fgh::Node::Map || over.kind != fgh::Node::Map) return over;
    for (const auto &f : over.fields) {
        bool found = false;
        for (auto &b : base.fields)
            if (b.first == f.first) {
                b.second = merge_settings(b.second, f.second);
                found = true;
                break;
            }
        if (!found) base.fields.push_back(f);
    }
    return base;
}

using fgh::Node;
using fgh::SettingsError;

constexpr double MU0 = 1.25663706127e-6;
const char *VERSION = "feellgood_amd 1.0 (gfx950)";

struct Config {
    std::string settings_name;
    std::string out_dir = ".", basename;
    double evol_dt = 5e-13, final_time = 1e-10;
    std::vector<std::string> cols = {"t", "<Mx>", "<My>", "<Mz>", "E_ex", "E_demag", "E_zeeman", "E_tot"};
    long mag_every = 100;  // 0: no .sol output
    std::vector<std::pair<std::string, std::string>> metadata;
    std::string mesh_file;
    double length_unit = 1e-9;
    std::vector<std::string> vol_names, surf_names;
    std::vector<fg_volume_region> vol;
    std::vector<fg_surface_region> surf;
    // spin accumulation (src/settings.cpp:414-418,438-449,545-550; defaults of
    // src/default-settings.yml: zero volume parameters, null surface ones)
    bool spin_acc = false, V_file = false;
    std::vector<fg_spin_volume> spin_vol;
    std::vector<fg_spin_surface> spin_surf;
    std::vector<double> surf_V;
    std::string init_sol;
    double m0[3] = {0, 0, 1};
    std::vector<fgh::Expr> m0_expr;  // three expressions of x, y, z (empty: m0 or init_sol)
    std::vector<std::string> m0_text;
    bool has_t0 = false;
    double t0 = 0;
    double Bext[3] = {0, 0, 0};
    std::vector<fgh::Expr> B_expr;  // three expressions of t (empty: Bext)
    std::vector<std::string> B_text;
    // R4toR3 field H = time(t) / mu0 * space(x) (src/settings.cpp:517-537): space of x, y, z at the
    // tetrahedra's Gauss points, time of t
    std::vector<fgh::Expr> Bs_expr, Bt_expr;
    std::vector<std::string> Bs_text, Bt_text;
    long max_iter = 700;
    double tol = 1e-6;
    double du_max = 0.02, dt_min = 1e-16, dt_max = 5e-13;
    int demag = FG_DEMAG_FMM, fmm_order = 8, fmm_leaf = 64, precision = FG_PREC_FP64, device = 0;
    int krylov = FG_KRYLOV_FP64;
};

// Three expressions of a settings key (numbers are expressions too). Returns whether any depends
// on the variables; if none does, their values are in val.
bool compile3(const Node &n, const std::vector<std::string> &vars, const std::string &where, double val[3],
              std::vector<fgh::Expr> &ex, std::vector<std::string> &text)
{
    if (n.kind != Node::Seq || n.items.size() != 3) throw SettingsError(where + " should have three components");
    ex.clear();
    text.clear();
    bool var = false;
    for (int k = 0; k < 3; k++) {
        const Node &it = n.items[k];
        if (it.kind != Node::Scalar || it.is_null())
            throw SettingsError(where + ": component " + std::to_string(k + 1) + " is not an expression");
        try {
            ex.push_back(fgh::Expr::compile(it.text, vars));
        } catch (const fgh::ExprError &e) {
            throw SettingsError(where + ": " + e.what());
        }
        text.push_back(it.text);
        var = var || ex.back().uses_variables();
    }
    if (!var)
        for (int k = 0; k < 3; k++) val[k] = ex[k].eval(nullptr);
    return var;
}

// A JavaScript function definition in a settings key, "function(a, b, ...) { return [e1, e2, e3]; }"
// (ExpressionParser::set_function, src/expression_parser.cpp:60-71), limited to a body that is one
// return statement of a three-element array (ncomp 3) or of one expression (ncomp 1): compiled as
// expressions of its parameters, whose number must be nparams. Returns whether any component
// depends on them (else the values are in val).
bool compile_function(const std::string &src, size_t nparams, const std::string &where, double val[3],
                      std::vector<fgh::Expr> &ex, std::vector<std::string> &text, size_t ncomp = 3)
{
    auto bad = [&](const std::string &why) {
        return SettingsError(where + ": " + why + " (supported: function(...) { return [e1, e2, e3]; })");
    };
    auto trim = [](std::string t) {
        const size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
    };
    const std::string f = trim(src);
    if (f.rfind("function", 0) != 0) throw bad("not a function definition");
    const size_t lp = f.find('('), rp = f.find(')', lp == std::string::npos ? 0 : lp);
    const size_t lb = f.find('{', rp == std::string::npos ? 0 : rp), rb = f.rfind('}');
    if (lp == std::string::npos || rp == std::string::npos || lb == std::string::npos || rb == std::string::npos ||
        rb < lb || !trim(f.substr(rb + 1)).empty())
        throw bad("malformed function");
    std::vector<std::string> params;
    std::stringstream ps(f.substr(lp + 1, rp - lp - 1));
    for (std::string q; std::getline(ps, q, ',');)
        if (!trim(q).empty()) params.push_back(trim(q));
    if (params.size() != nparams)
        throw bad("the function should take " + std::to_string(nparams) + " parameter(s)");
    std::string body = trim(f.substr(lb + 1, rb - lb - 1));
    if (body.rfind("return", 0) != 0) throw bad("the body should be one return statement");
    body = trim(body.substr(6));
    if (!body.empty() && body.back() == ';') body = trim(body.substr(0, body.size() - 1));
    std::vector<std::string> comp(1);
    if (ncomp == 1) {
        comp[0] = body;
    } else {
        if (body.size() < 2 || body.front() != '[' || body.back() != ']')
            throw bad("the function should return [x, y, z]");
        // the three top-level components of the array
        int depth = 0;
        for (size_t k = 1; k + 1 < body.size(); k++) {
            const char ch = body[k];
            depth += (ch == '(' || ch == '[') - (ch == ')' || ch == ']');
            if (ch == ',' && depth == 0)
                comp.emplace_back();
            else
                comp.back() += ch;
        }
        if (comp.size() != 3) throw bad("the function should return three components");
    }
    ex.clear();
    text.clear();
    bool var = false;
    for (auto &cpt : comp) {
        try {
            ex.push_back(fgh::Expr::compile(trim(cpt), params));
        } catch (const fgh::ExprError &e) {
            throw SettingsError(where + ": " + e.what());
        }
        text.push_back(trim(cpt));
        var = var || ex.back().uses_variables();
    }
    if (!var)
        for (size_t k = 0; k < ncomp; k++) val[k] = ex[k].eval(nullptr);
    return var;
}

void normalize(double v[3], const std::string &where)
{
    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n == 0) throw SettingsError(where + " should not be the null vector");
    for (int k = 0; k < 3; k++) v[k] /= n;
}

fg_volume_region default_volume()
{
    // src/default-settings.yml: mesh.volume_regions.__default__
    fg_volume_region r{};
    r.alpha = 0.5;
    r.Ae = 1e-11;
    r.Ms = 795774.7;
    r.uk[2] = 1;
    r.ex[0] = 1;
    r.ey[1] = 1;
    r.ez[2] = 1;
    return r;
}

void check_map(const Node *n, const std::string &where)
{
    if (n && !n->is_null() && n->kind != Node::Map) throw SettingsError(where + " should be a map");
}

Config read_config(const Node &root)
{
    Config c;
    if (root.kind != Node::Map) throw SettingsError("the settings should be a map");
    const Node *o = root.get("outputs");
    check_map(o, "outputs");
    if (o && o->kind == Node::Map) {
        if (auto n = o->get("directory"); n && !n->is_null()) {
            c.out_dir = fgh::as_string(*n, "outputs.directory");
            if (c.out_dir.empty()) c.out_dir = ".";
            if (c.out_dir.size() > 1 && c.out_dir.back() == '/') c.out_dir.pop_back();
        }
        if (auto n = o->get("file_basename"); n && !n->is_null()) c.basename = fgh::as_string(*n, "outputs.file_basename");
        if (auto n = o->get("evol_time_step"); n && !n->is_null()) c.evol_dt = fgh::as_double(*n, "outputs.evol_time_step");
        if (auto n = o->get("final_time"); n && !n->is_null()) c.final_time = fgh::as_double(*n, "outputs.final_time");
        if (auto n = o->get("evol_columns"); n && !n->is_null()) {
            if (n->kind != Node::Seq || n->items.empty()) throw SettingsError("outputs.evol_columns should be a list");
            c.cols.clear();
            for (auto &it : n->items) c.cols.push_back(fgh::as_string(it, "outputs.evol_columns"));
        }
        if (auto n = o->get("mag_config_every"); n && !n->is_null()) {
            if (n->kind == Node::Scalar && (n->text == "false" || n->text == "False")) c.mag_every = 0;
            else c.mag_every = fgh::as_long(*n, "outputs.mag_config_every");
        }
        if (auto n = o->get("metadata"); n && !n->is_null()) {
            if (n->kind != Node::Map) throw SettingsError("outputs.metadata should be a map.");
            for (auto &f : n->fields) c.metadata.emplace_back(f.first, f.second.is_null() ? "" : f.second.text);
        }
    }
    const Node *m = root.get("mesh");
    if (!m || m->kind != Node::Map) throw SettingsError("mesh.filename is required");
    if (auto n = m->get("filename"); n && !n->is_null()) c.mesh_file = fgh::as_string(*n, "mesh.filename");
    if (c.mesh_file.empty()) throw SettingsError("mesh.filename is required");
    if (auto n = m->get("length_unit"); n && !n->is_null()) c.length_unit = fgh::as_double(*n, "mesh.length_unit");
    const Node *vr = m->get("volume_regions");
    check_map(vr, "mesh.volume_regions");
    if (vr && vr->kind == Node::Map)
        for (auto &f : vr->fields) {
            const std::string w = "mesh.volume_regions." + f.first;
            check_map(&f.second, w);
            fg_volume_region r = default_volume();
            fg_spin_volume sv{0.0, 0.0, 0.0, 0.0, 0.0};
            const Node &p = f.second;
            if (p.kind == Node::Map) {
                if (auto n = p.get("sigma")) sv.sigma = fgh::as_double(*n, w + ".sigma");
                if (auto n = p.get("dens_state")) sv.N0 = fgh::as_double(*n, w + ".dens_state");
                if (auto n = p.get("P")) sv.P = fgh::as_double(*n, w + ".P");
                if (auto n = p.get("l_sd")) sv.lsd = fgh::as_double(*n, w + ".l_sd");
                if (auto n = p.get("l_sf")) sv.lsf = fgh::as_double(*n, w + ".l_sf");
                if (auto n = p.get("Ae")) r.Ae = fgh::as_double(*n, w + ".Ae");
                if (auto n = p.get("Ms")) r.Ms = fgh::as_double(*n, w + ".Ms");
                if (auto n = p.get("K")) r.K = fgh::as_double(*n, w + ".K");
                if (auto n = p.get("uk")) fgh::as_vec3(*n, r.uk, w + ".uk");
                if (auto n = p.get("K3")) r.K3 = fgh::as_double(*n, w + ".K3");
                if (auto n = p.get("ex")) fgh::as_vec3(*n, r.ex, w + ".ex");
                if (auto n = p.get("ey")) fgh::as_vec3(*n, r.ey, w + ".ey");
                if (auto n = p.get("ez")) fgh::as_vec3(*n, r.ez, w + ".ez");
                if (auto n = p.get("alpha_LLG")) r.alpha = fgh::as_double(*n, w + ".alpha_LLG");
            }
            normalize(r.uk, w + ".uk");
            normalize(r.ex, w + ".ex");
            normalize(r.ey, w + ".ey");
            normalize(r.ez, w + ".ez");
            c.vol_names.push_back(f.first);
            c.vol.push_back(r);
            c.spin_vol.push_back(sv);
        }
    if (c.vol.empty()) throw SettingsError("mesh.volume_regions should name at least one region");
    const Node *sr = m->get("surface_regions");
    check_map(sr, "mesh.surface_regions");
    if (sr && sr->kind == Node::Map)
        for (auto &f : sr->fields) {
            const std::string w = "mesh.surface_regions." + f.first;
            check_map(&f.second, w);
            fg_surface_region r{};
            r.uk[2] = 1;
            fg_spin_surface ss{NAN, {NAN, NAN, NAN}, {NAN, NAN, NAN}};
            double V = NAN;
            const Node &p = f.second;
            if (p.kind == Node::Map) {
                // jn and V may be null, read as NaN (src/settings.cpp:438-449)
                if (auto n = p.get("jn"); n && !n->is_null()) ss.jn = fgh::as_double(*n, w + ".jn");
                if (auto n = p.get("V"); n && !n->is_null()) V = fgh::as_double(*n, w + ".V");
                if (!std::isnan(ss.jn) && !std::isnan(V))
                    throw SettingsError(w + ": a surface region cannot have both no-null J and V");
                if (auto n = p.get("uP"); n && !n->is_null()) {
                    fgh::as_vec3(*n, ss.uP, w + ".uP");
                    normalize(ss.uP, w + ".uP");
                }
                if (auto n = p.get("s"); n && !n->is_null()) fgh::as_vec3(*n, ss.s, w + ".s");
                if (auto n = p.get("Ks")) r.Ks = fgh::as_double(*n, w + ".Ks");
                if (auto n = p.get("uk")) fgh::as_vec3(*n, r.uk, w + ".uk");
                if (auto n = p.get("suppress_charges")) r.suppress_charges = fgh::as_bool(*n, w + ".suppress_charges");
            }
            normalize(r.uk, w + ".uk");
            c.surf_names.push_back(f.first);
            c.surf.push_back(r);
            c.spin_surf.push_back(ss);
            c.surf_V.push_back(V);
        }
    if (auto n = root.get("initial_magnetization"); n && !n->is_null()) {
        if (n->kind == Node::Scalar) {
            // src/settings.cpp:457-466: a function definition, else a .sol file name
            if (n->text.find("function") != std::string::npos && n->text.find('{') != std::string::npos) {
                if (!compile_function(n->text, 3, "initial_magnetization", c.m0, c.m0_expr, c.m0_text))
                    c.m0_expr.clear();
            } else {
                c.init_sol = n->text;
            }
        } else if (!compile3(*n, {"x", "y", "z"}, "initial_magnetization", c.m0, c.m0_expr, c.m0_text)) {
            c.m0_expr.clear();  // constants: c.m0
        }
    }
    if (auto n = root.get("initial_time"); n && !n->is_null()) {
        c.has_t0 = true;
        c.t0 = fgh::as_double(*n, "initial_time");
    }
    if (auto n = root.get("spin_accumulation"); n && !n->is_null()) {
        check_map(n, "spin_accumulation");
        if (auto b = n->get("enable")) c.spin_acc = fgh::as_bool(*b, "spin_accumulation.enable");
        if (auto b = n->get("V_file")) c.V_file = fgh::as_bool(*b, "spin_accumulation.V_file");
    }
    if (auto n = root.get("Bext"); n && !n->is_null()) {
        if (n->kind == Node::Scalar) {
            if (!compile_function(n->text, 1, "Bext", c.Bext, c.B_expr, c.B_text)) c.B_expr.clear();
        } else if (n->kind == Node::Seq) {
            if (!compile3(*n, {"t"}, "Bext", c.Bext, c.B_expr, c.B_text)) c.B_expr.clear();
        } else {
            const Node *sp = n->get("space"), *tm = n->get("time");
            if (!sp || !tm || sp->kind != Node::Scalar || tm->kind != Node::Scalar)
                throw SettingsError("Bext should be a function, three expressions, or space and time functions");
            double dummy[3];
            compile_function(sp->text, 3, "Bext.space", dummy, c.Bs_expr, c.Bs_text);
            compile_function(tm->text, 1, "Bext.time", dummy, c.Bt_expr, c.Bt_text, 1);
        }
    }
    if (auto r = root.get("recentering"); r && r->kind == Node::Map)
        if (auto n = r->get("enable"); n && !n->is_null() && fgh::as_bool(*n, "recentering.enable"))
            throw SettingsError("recentering is not supported");
    if (auto s = root.get("finite_element_solver"); s && s->kind == Node::Map) {
        if (auto n = s->get("max(iter)"); n && !n->is_null()) c.max_iter = fgh::as_long(*n, "finite_element_solver.max(iter)");
        if (auto n = s->get("tolerance"); n && !n->is_null()) c.tol = fgh::as_double(*n, "finite_element_solver.tolerance");
    }
    if (auto s = root.get("time_integration"); s && s->kind == Node::Map) {
        if (auto n = s->get("max(du)"); n && !n->is_null()) c.du_max = fgh::as_double(*n, "time_integration.max(du)");
        if (auto n = s->get("min(dt)"); n && !n->is_null()) c.dt_min = fgh::as_double(*n, "time_integration.min(dt)");
        if (auto n = s->get("max(dt)"); n && !n->is_null()) c.dt_max = fgh::as_double(*n, "time_integration.max(dt)");
    }
    // engine section (not in the reference): demag method, FMM order and leaf size, precision
    if (auto s = root.get("gpu"); s && s->kind == Node::Map) {
        if (auto n = s->get("demag"); n && !n->is_null()) {
            const std::string v = fgh::as_string(*n, "gpu.demag");
            if (v == "fmm") c.demag = FG_DEMAG_FMM;
            else if (v == "direct") c.demag = FG_DEMAG_DIRECT;
            else throw SettingsError("gpu.demag should be fmm or direct");
        }
        if (auto n = s->get("fmm_order"); n && !n->is_null()) c.fmm_order = (int)fgh::as_long(*n, "gpu.fmm_order");
        if (auto n = s->get("fmm_leaf"); n && !n->is_null()) c.fmm_leaf = (int)fgh::as_long(*n, "gpu.fmm_leaf");
        if (auto n = s->get("precision"); n && !n->is_null()) {
            const std::string v = fgh::as_string(*n, "gpu.precision");
            if (v == "fp64") c.precision = FG_PREC_FP64;
            else if (v == "mixed") c.precision = FG_PREC_MIXED;
            else throw SettingsError("gpu.precision should be fp64 or mixed");
        }
        if (auto n = s->get("krylov"); n && !n->is_null()) {
            const std::string v = fgh::as_string(*n, "gpu.krylov");
            if (v == "fp64") c.krylov = FG_KRYLOV_FP64;
            else if (v == "mixed") c.krylov = FG_KRYLOV_MIXED;
            else throw SettingsError("gpu.krylov should be fp64 or mixed");
        }
        if (auto n = s->get("device"); n && !n->is_null()) c.device = (int)fgh::as_long(*n, "gpu.device");
    }
    for (auto &col : c.cols) {
        static const char *known[] = {"iter", "t", "dt", "max_dm", "max_angle", "<Mx>", "<My>", "<Mz>", "<dMx/dt>",
                                      "<dMy/dt>", "<dMz/dt>", "E_ex", "E_aniso", "E_demag", "E_zeeman", "E_tot",
                                      "Hx", "Hy", "Hz"};
        if (std::find_if(std::begin(known), std::end(known), [&](const char *k) { return col == k; }) == std::end(known))
            throw SettingsError("outputs.evol_columns: unsupported column '" + col + "'");
    }
    if (c.basename.empty()) {
        std::string b = c.mesh_file;
        const size_t sl = b.rfind('/');
        if (sl != std::string::npos) b = b.substr(sl + 1);
        const size_t dot = b.rfind('.');
        if (dot != std::string::npos && dot > 0) b = b.substr(0, dot);
        c.basename = b;
    }
    return c;
}

std::string date_string()
{
    const std::time_t t = std::time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%S%z", std::localtime(&t));
    return buf;
}

std::string common_metadata(const Config &c)
{
    std::ostringstream ss;
    ss << "## feeLLGood version: " << VERSION << "\n";
    char name[HOST_NAME_MAX + 1];
    if (gethostname(name, sizeof name) == 0) ss << "## hostname: " << name << "\n";
    ss << "## real-world time: " << date_string() << "\n";
    ss << "## settings file: " << c.settings_name << "\n";
    for (auto &kv : c.metadata) {
        std::string v = kv.second;
        if (!v.empty() && v.back() == '\n') v.pop_back();
        std::string esc;
        for (char ch : v) esc += ch == '\n' ? std::string("\n##  ") : std::string(1, ch);
        ss << "## " << kv.first << ": " << esc << "\n";
    }
    return ss.str();
}

struct Engine {
    fg_ctx *ctx = nullptr;
    void check(int rc, const char *what)
    {
        if (rc < 0) {
            std::ostringstream o;
            o << what << ": " << fg_last_error(ctx);
            throw std::runtime_error(o.str());
        }
    }
};

// Settings::getMagnetization (src/settings.h:262-268): the three expressions at the node position
// (metres), normalised.
void initial_from_expressions(const Config &c, const fg_mesh &view, double *u)
{
    for (int i = 0; i < view.n_nodes; i++) {
        double v[3];
        for (int k = 0; k < 3; k++) v[k] = c.m0_expr[k].eval(view.pos + 3 * (size_t)i);
        normalize(v, "initial_magnetization at node " + std::to_string(i));
        for (int k = 0; k < 3; k++) u[3 * (size_t)i + k] = v[k];
    }
}

int run(const Config &c, unsigned seed, bool verbose, bool check_only, const std::string &steps_log)
{
    char err[512] = {0};
    fgh_mesh *mesh = fgh_read_msh(c.mesh_file.c_str(), err, sizeof err);
    if (!mesh) throw std::runtime_error(std::string("mesh: ") + err);
    std::vector<const char *> vn, sn;
    for (auto &s : c.vol_names) vn.push_back(s.c_str());
    for (auto &s : c.surf_names) sn.push_back(s.c_str());
    if (fgh_prepare(mesh, c.length_unit, (int)vn.size(), vn.data(), c.vol.data(), (int)sn.size(), sn.data(),
                    c.surf.data()) != FG_OK)
        throw std::runtime_error(std::string("mesh: ") + fgh_last_error());
    fg_mesh view;
    fgh_view(mesh, &view);
    const int N = view.n_nodes;
    std::cout << "mesh:              " << c.mesh_file << " (" << N << " nodes, " << view.n_tets << " tetrahedra, "
              << view.n_tris << " triangles)\n";
    if (check_only) {
        // settings and mesh only, no device: what the run would use
        std::cout.precision(10);
        for (size_t k = 0; k < c.vol.size(); k++)
            std::cout << "volume region:     " << c.vol_names[k] << " Ae " << c.vol[k].Ae << " Ms " << c.vol[k].Ms
                      << " K " << c.vol[k].K << " uk " << c.vol[k].uk[0] << " " << c.vol[k].uk[1] << " "
                      << c.vol[k].uk[2] << " K3 " << c.vol[k].K3 << " alpha " << c.vol[k].alpha << "\n";
        for (size_t k = 0; k < c.surf.size(); k++)
            std::cout << "surface region:    " << c.surf_names[k] << " Ks " << c.surf[k].Ks << " suppress_charges "
                      << c.surf[k].suppress_charges << "\n";
        if (c.spin_acc) {
            std::cout << "spin accumulation: V_file " << c.V_file << "\n";
            for (size_t k = 0; k < c.vol.size(); k++)
                std::cout << "spin volume:       " << c.vol_names[k] << " sigma " << c.spin_vol[k].sigma << " N0 "
                          << c.spin_vol[k].N0 << " P " << c.spin_vol[k].P << " lsd " << c.spin_vol[k].lsd << " lsf "
                          << c.spin_vol[k].lsf << "\n";
            for (size_t k = 0; k < c.surf.size(); k++)
                std::cout << "spin surface:      " << c.surf_names[k] << " jn " << c.spin_surf[k].jn << " V "
                          << c.surf_V[k] << " uP " << c.spin_surf[k].uP[0] << " " << c.spin_surf[k].uP[1] << " "
                          << c.spin_surf[k].uP[2] << " s " << c.spin_surf[k].s[0] << " " << c.spin_surf[k].s[1]
                          << " " << c.spin_surf[k].s[2] << "\n";
        }
        if (!c.Bs_expr.empty())
            std::cout << "Bext(x,t):         space " << c.Bs_text[0] << " | " << c.Bs_text[1] << " | " << c.Bs_text[2]
                      << " time " << c.Bt_text[0] << "\n";
        else if (c.B_expr.empty())
            std::cout << "Bext:              " << c.Bext[0] << " " << c.Bext[1] << " " << c.Bext[2] << "\n";
        else
            std::cout << "Bext(t):           " << c.B_text[0] << " | " << c.B_text[1] << " | " << c.B_text[2] << "\n";
        if (!c.m0_expr.empty()) {
            // the normalised initial magnetization of the nodes, averaged
            std::vector<double> u0(3 * (size_t)N);
            initial_from_expressions(c, view, u0.data());
            double a[3] = {0, 0, 0};
            for (int i = 0; i < N; i++)
                for (int k = 0; k < 3; k++) a[k] += u0[3 * (size_t)i + k] / N;
            std::cout << "initial:           expressions " << c.m0_text[0] << " | " << c.m0_text[1] << " | "
                      << c.m0_text[2] << " mean " << a[0] << " " << a[1] << " " << a[2] << "\n";
        } else {
            std::cout << "initial:           " << (c.init_sol.empty() ? "constant" : c.init_sol) << " " << c.m0[0]
                      << " " << c.m0[1] << " " << c.m0[2] << "\n";
        }
        std::cout << "time:              final " << c.final_time << " evol " << c.evol_dt << " dt " << c.dt_min << " "
                  << c.dt_max << " du " << c.du_max << "\n";
        std::cout << "solver:            tol " << c.tol << " max_iter " << c.max_iter << "\n";
        std::cout << "outputs:           " << c.out_dir << "/" << c.basename << " every " << c.mag_every << " columns";
        for (auto &col : c.cols) std::cout << " " << col;
        std::cout << "\n";
        fgh_free(mesh);
        return 0;
    }

    std::vector<double> u(3 * (size_t)N), phi(N, 0.0), v(3 * (size_t)N, 0.0);
    double t = c.has_t0 ? c.t0 : 0.0;
    if (!c.init_sol.empty()) {
        double ts = 0;
        if (fgh_read_sol(mesh, c.init_sol.c_str(), u.data(), phi.data(), &ts) != FG_OK)
            throw std::runtime_error(std::string("initial magnetization: ") + fgh_last_error());
        if (!c.has_t0) t = ts;
    } else if (!c.m0_expr.empty()) {
        initial_from_expressions(c, view, u.data());
    } else {
        double m0[3] = {c.m0[0], c.m0[1], c.m0[2]};
        normalize(m0, "initial_magnetization");
        for (int i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) u[3 * (size_t)i + k] = m0[k];
    }

    Engine e;
    fg_options opt;
    fg_default_options(&opt);
    opt.tol = c.tol;
    opt.max_iter = (int)c.max_iter;
    opt.demag_method = c.demag;
    opt.fmm_order = c.fmm_order;
    opt.fmm_leaf = c.fmm_leaf;
    opt.precision = c.precision;
    opt.krylov = c.krylov;
    opt.device = c.device;
    if (fg_create(&view, &opt, &e.ctx) != FG_OK) throw std::runtime_error(std::string("fg_create: ") + fg_last_error(nullptr));

    // src/main.cpp:211 and the per-step draw of src/linear_algebra.cpp:5-10
    srand(seed);
    auto next_angle = []() {
        std::mt19937 gen(rand());
        std::uniform_real_distribution<> distrib(0.0, 1.0);
        return M_2_PI * distrib(gen);
    };
    e.check(fg_set_state(e.ctx, FG_BOTH, u.data(), nullptr, phi.data(), nullptr), "fg_set_state");
    e.check(fg_base_projection(e.ctx, next_angle()), "fg_base_projection");  // LinAlgebra ctor

    // Settings::getField (src/settings.h:275-283): H = B(t) / mu0; with a space-dependent field
    // (getFieldSpace, getFieldTime, :309-317) the space part at every Gauss point goes to the
    // engine once (src/mesh.cpp:377-389) and each step passes the amplitude time(t) / mu0
    const bool spacef = !c.Bs_expr.empty();
    double Hext[3], amp = 0;
    auto field_at = [&](double tt) {
        if (spacef) {
            amp = c.Bt_expr[0].eval(&tt) / MU0;
            for (int k = 0; k < 3; k++) Hext[k] = NAN;  // no uniform field to report
        } else {
            for (int k = 0; k < 3; k++) Hext[k] = (c.B_expr.empty() ? c.Bext[k] : c.B_expr[k].eval(&tt)) / MU0;
        }
    };
    if (spacef) {
        std::vector<double> pg(15 * (size_t)view.n_tets), h(pg.size());
        e.check(fg_gauss_points(&view, pg.data()), "fg_gauss_points");
        for (size_t q = 0; q < pg.size(); q += 3)
            for (int k = 0; k < 3; k++) h[q + k] = c.Bs_expr[k].eval(&pg[q]);
        e.check(fg_set_space_field(e.ctx, h.data()), "fg_set_space_field");
    }
    field_at(t);
    // timing (src/time_integration.h:6-59)
    const double DTMIN = c.dt_min, DTMAX = c.dt_max, TAUR = 100. * DTMAX;
    double dt = 0, prefactor = 1;
    auto set_dt = [&](double d) {
        dt = d;
        const double tt = d / TAUR;
        prefactor = 1. + tt * std::fabs(std::log(tt));
    };
    set_dt(std::sqrt(DTMIN * DTMAX));
    double E[4] = {0, 0, 0, 0};
    double vmax = 0;
    // spin accumulation: the potential once, then s from the magnetization at every compute_all
    // (spinAcc::spinAcc and compute, src/spinAccumulationSolver.cpp:8-36,151-163); the exchange
    // reads the CURRENT state and the source the NEXT one, as the reference does
    std::vector<double> Vpot, sacc;
    // per region of the mesh view: the settings' regions first, then the surfaces the preparation
    // adds (controlTriangles), which set neither jn, V, uP nor s
    std::vector<fg_spin_volume> svol(c.spin_vol);
    svol.resize(std::max<size_t>(svol.size(), (size_t)view.n_vol_regions), fg_spin_volume{0.0, 0.0, 0.0, 0.0, 0.0});
    std::vector<fg_spin_surface> ssurf(c.spin_surf);
    std::vector<double> sV(c.surf_V);
    ssurf.resize(std::max<size_t>(ssurf.size(), (size_t)view.n_surf_regions),
                 fg_spin_surface{NAN, {NAN, NAN, NAN}, {NAN, NAN, NAN}});
    sV.resize(ssurf.size(), NAN);
    // the persistent solver (spinAcc, built once): every call adds its assembly to the system of
    // the earlier ones, as the reference's solver does (src/solver.h:110-142), and installs s, or
    // zero when it did not converge, as the field of the next element pass
    std::unique_ptr<fg_spin, int (*)(fg_spin *)> spin(nullptr, fg_spin_destroy);
    auto spin_compute = [&]() {
        fg_solve_info si;
        e.check(fg_spin_solve_ctx(spin.get(), e.ctx, 1, sacc.data(), &si), "fg_spin_solve_ctx");
        return si.status == FG_ITER_CONVERGED;
    };
    if (c.spin_acc) {
        std::vector<double> sig, jn;
        for (auto &p : svol) sig.push_back(p.sigma);
        for (auto &p : ssurf) jn.push_back(p.jn);
        Vpot.resize(N);
        sacc.resize(3 * (size_t)N);
        fg_solve_info si;
        const int rc = fg_electrostatic_potential(&view, sig.data(), jn.data(), sV.data(), 1e-8, 1000, c.device,
                                                  Vpot.data(), &si);
        if (rc == FG_ERR_ARG) throw std::runtime_error("incorrect boundary conditions for potential V solver");
        e.check(rc, "fg_electrostatic_potential");
        if (si.status != FG_ITER_CONVERGED) throw std::runtime_error("electrostatic solver did not converge");
        if (c.V_file) {  // electrostatSolver::save (src/electrostatSolver.cpp:138-156), in the working
                         // directory as the reference names it (src/spinAccumulationSolver.cpp:26-27)
            const std::string vname = c.basename + "_V.sol";
            std::ofstream vf(vname);
            if (!vf) throw std::runtime_error("cannot open " + vname);
            vf << "## real-world time: " << date_string() << "\n## columns: index\tV\n" << std::scientific
               << std::setprecision(8);
            const int32_t *ni = fgh_node_index(mesh);
            for (int i = 0; i < N; i++) vf << i << '\t' << Vpot[ni[i]] << "\n";
        }
        fg_spin *sp = nullptr;
        const int src = fg_spin_create(&view, svol.data(), ssurf.data(), Vpot.data(), 1e-8, 1000, c.device, &sp);
        if (src == FG_ERR_ARG) throw std::runtime_error("incorrect boundary conditions for spin diffusion solver");
        e.check(src, "fg_spin_create");
        spin.reset(sp);
        if (!spin_compute()) throw std::runtime_error("spin diffusion solver (first try) failed");
    }
    auto compute_all = [&]() {  // src/fem.h:203-224
        e.check(fg_calc_demag(e.ctx), "fg_calc_demag");
        if (c.spin_acc && !spin_compute() && verbose) std::cerr << "spin diffusion solver failed.\n";
        e.check(spacef ? fg_energy_amp(e.ctx, amp, E) : fg_energy(e.ctx, Hext, E), "fg_energy");
        e.check(fg_evolution(e.ctx), "fg_evolution");
    };
    compute_all();

    mkdir(c.out_dir.c_str(), 0777);
    const std::string base = c.out_dir + "/" + c.basename;
    std::ofstream fout(base + ".evol");
    if (!fout) throw std::runtime_error("cannot open " + base + ".evol");
    fout << common_metadata(c) << "## columns: ";
    for (size_t k = 0; k < c.cols.size(); k++) fout << c.cols[k] << (k + 1 < c.cols.size() ? "\t" : "\n");
    fout.precision(16);

    const double t_initial = t, t_step = c.evol_dt;
    const int step_count = (int)std::lround((c.final_time - t_initial) / t_step);
    // TimeStepper (src/time_integration.cpp:11-38)
    const double hard_min = DTMIN, hard_max = DTMAX * (1 + FLT_EPSILON);
    double soft_max = dt;
    auto stepper = [&](double stride) {
        double step = std::min(stride, soft_max);
        if (step > stride - 2 * hard_min && step < stride) step = stride - 2 * hard_min;
        soft_max = std::max(soft_max, std::min(step * 1.1, hard_max));
        return step;
    };
    auto set_soft_limit = [&](double mx) { soft_max = std::min(soft_max, mx); };

    int status = 0, nt = 0, flag = 0, nt_output = 0, bad = 0;
    long total_iter = 0;
    double max_angle = 0;
    e.check(fg_max_angle(e.ctx, &max_angle), "fg_max_angle");
    const auto wall0 = std::chrono::steady_clock::now();
    // one line per step attempt: its start time, dt, timing prefactor, basis angle, applied field,
    // outcome (0 accepted, 1 solver failure, 2 rejected on max(du)) and BiCGSTAB iterations, so
    // that the same sequence of decisions can be replayed elsewhere
    std::ofstream slog;
    if (!steps_log.empty()) {
        slog.open(steps_log);
        if (!slog) throw std::runtime_error("cannot open " + steps_log);
        slog << "# t dt prefactor angle Hx Hy Hz outcome iterations\n" << std::setprecision(17);
    }
    auto log_attempt = [&](double r, int outcome, int iters) {
        if (slog)
            slog << t << ' ' << dt << ' ' << prefactor << ' ' << r << ' ' << Hext[0] << ' ' << Hext[1] << ' '
                 << Hext[2] << ' ' << outcome << ' ' << iters << '\n';
    };
    // state saved on SIGINT / SIGTERM (exit_if_signal_received, src/time_integration.cpp:99-134)
    auto exit_if_signal_received = [&]() {
        if (!received_signal) return;
        const char *name = received_signal == SIGINT ? "SIGINT" : received_signal == SIGTERM ? "SIGTERM" : "signal";
        std::cout << "\nReceived " << name;
        if (c.mag_every > 0) {
            std::cout << ": saving the magnetization configuration...\n";
            const std::string fname = base + "_at_exit.sol";
            e.check(fg_get_state(e.ctx, FG_NEXT, u.data(), nullptr, phi.data(), nullptr), "fg_get_state");
            std::ostringstream md;
            md << common_metadata(c) << "## time: " << std::scientific << t << "\n"
               << "## columns: idx\tmx\tmy\tmz\tphi" << (c.spin_acc ? "\tsx\tsy\tsz" : "") << "\n";
            if (fgh_write_sol_s(mesh, fname.c_str(), md.str().c_str(), 7, u.data(), phi.data(),
                                c.spin_acc ? sacc.data() : nullptr) != FG_OK)
                throw std::runtime_error(std::string("writing ") + fname + ": " + fgh_last_error());
            std::cout << "Magnetization configuration saved to " << fname << "\n";
        } else {
            std::cout << ": magnetization configuration not saved.\n";
        }
        std::cout << "Terminating.\n";
        std::cout << "steps:             " << nt << " accepted, " << bad << " rejected, " << total_iter
                  << " BiCGSTAB iterations\n" << std::flush;
        fout.close();
        slog.close();
        const int sig = received_signal;
        fg_destroy(e.ctx);
        fgh_free(mesh);
        signal(sig, SIG_DFL);  // the process ends with the signal's own status
        raise(sig);
        exit(1);
    };
    for (int step_nb = 0; step_nb <= step_count; step_nb++) {
        const double t_target = t_initial + step_nb * t_step;
        while (t < t_target) {
            exit_if_signal_received();
            set_dt(stepper(t_target - t));
            const bool last_step = (dt == t_target - t);
            if (verbose)
                std::cout << "evol step = " << nt_output << ", step = " << nt << ", t = " << t << ", dt = " << dt
                          << (flag ? "  (retry)" : "") << '\n';
            if (dt < DTMIN) {
                std::cout << "\n**ABORTED**: dt < DTMIN\n";
                status = 1;
                goto bailout;
            }
            const double r_attempt = next_angle();
            field_at(t);  // the field at the step's start, also used by its energies (src/time_integration.cpp:194-231)
            // a uniform field without spin accumulation: the loop body in one call (fg_step), one
            // host read at its end; otherwise the calls one by one
            const bool one_call = !spacef && !c.spin_acc && c.du_max > 0;  // fg_step's dumax test is this loop's
            fg_solve_info info;
            double E_step[4] = {0, 0, 0, 0}, a_step = 0;
            int rc;
            if (one_call) {
                rc = fg_step(e.ctx, Hext, dt, prefactor, r_attempt, c.du_max, &info, E_step, &a_step);
                e.check(rc, "fg_step");
            } else {
                e.check(fg_base_projection(e.ctx, r_attempt), "fg_base_projection");
                e.check(spacef ? fg_prepare_elements_amp(e.ctx, amp, dt, prefactor)
                               : fg_prepare_elements(e.ctx, Hext, dt, prefactor),
                        "fg_prepare_elements");
                rc = fg_solve(e.ctx, dt, &info);
                e.check(rc, "fg_solve");
            }
            vmax = info.v_max;
            total_iter += info.iterations;
            if (rc == FG_SOLVE_FAILED) {
                log_attempt(r_attempt, 1, info.iterations);
                flag++;
                bad++;
                set_soft_limit(dt / 2);
                continue;
            }
            const double dumax = dt * vmax;
            set_soft_limit((c.du_max / vmax) * 0.95);
            if (dumax > c.du_max) {
                log_attempt(r_attempt, 2, info.iterations);
                flag++;
                bad++;
                continue;
            }
            log_attempt(r_attempt, 0, info.iterations);
            double a = a_step;
            if (one_call) {
                for (int k = 0; k < 4; k++) E[k] = E_step[k];
            } else {
                compute_all();
                e.check(fg_max_angle(e.ctx, &a), "fg_max_angle");
            }
            nt++;
            flag = 0;
            if (last_step) t = t_target;
            else t += dt;
            max_angle = std::max(max_angle, a);
        }
        field_at(t);  // the saved applied field is the one at the output time (src/save.cpp:22-24)
        {  // Fem::saver (src/save.cpp:13-162)
            double mu[3] = {0, 0, 0}, mv[3] = {0, 0, 0}, ang = 0;
            e.check(fg_average(e.ctx, 0, mu), "fg_average");
            for (size_t k = 0; k < c.cols.size(); k++) {
                const std::string &col = c.cols[k];
                if (col == "iter") fout << nt_output;
                else if (col == "t") fout << t;
                else if (col == "dt") fout << dt;
                else if (col == "max_dm") fout << vmax * dt;
                else if (col == "max_angle") {
                    e.check(fg_max_angle(e.ctx, &ang), "fg_max_angle");
                    fout << ang;
                } else if (col == "<Mx>") fout << mu[0];
                else if (col == "<My>") fout << mu[1];
                else if (col == "<Mz>") fout << mu[2];
                else if (col[0] == '<') {
                    e.check(fg_average(e.ctx, 1, mv), "fg_average");
                    fout << mv[col[3] - 'x'];
                } else if (col == "E_ex") fout << E[FG_E_EXCHANGE];
                else if (col == "E_aniso") fout << E[FG_E_ANISOTROPY];
                else if (col == "E_demag") fout << E[FG_E_DEMAG];
                else if (col == "E_zeeman") fout << E[FG_E_ZEEMAN];
                else if (col == "E_tot") fout << E[0] + E[1] + E[2] + E[3];
                else if (col == "Hx") fout << Hext[0];
                else if (col == "Hy") fout << Hext[1];
                else if (col == "Hz") fout << Hext[2];
                fout << (k + 1 < c.cols.size() ? "\t" : "\n");
            }
            fout << std::flush;
            if (c.mag_every && (nt_output % c.mag_every) == 0) {
                const std::string name = base + "_iter" + std::to_string(nt_output) + ".sol";
                e.check(fg_get_state(e.ctx, FG_NEXT, u.data(), nullptr, phi.data(), nullptr), "fg_get_state");
                std::ostringstream md;
                md << common_metadata(c) << "## time: " << std::scientific << t << "\n"
                   << "## columns: idx\tmx\tmy\tmz\tphi" << (c.spin_acc ? "\tsx\tsy\tsz" : "") << "\n";
                if (fgh_write_sol_s(mesh, name.c_str(), md.str().c_str(), 7, u.data(), phi.data(),
                                    c.spin_acc ? sacc.data() : nullptr) != FG_OK)
                    throw std::runtime_error(std::string("writing ") + name + ": " + fgh_last_error());
                if (verbose) std::cout << " " << name << "\n";
            }
            nt_output++;
        }
    }
bailout:
    fout.close();
    {
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
        std::cout << "steps:             " << nt << " accepted, " << bad << " rejected, " << total_iter
                  << " BiCGSTAB iterations\n";
        std::cout << "max angle:         " << max_angle << " rad\n";
        std::cout << "wall time:         " << wall << " s (" << (nt > 0 ? wall / nt * 1e3 : 0.0) << " ms per step)\n";
    }
    fg_destroy(e.ctx);
    fgh_free(mesh);
    return status;
}

void usage(const char *prog)
{
    std::cout << "Usage: " << prog << " [options] settings_file\n"
              << "  settings_file    YAML or JSON settings; '-' reads standard input\n"
              << "  -v, --verbose    print every integration step\n"
              << "  --seed N         fixed random seed (default: random)\n"
              << "  --check          read the settings and the mesh, print them, run nothing\n"
              << "  --steps-log FILE write every step attempt (t, dt, prefactor, basis angle, field,\n"
              << "                   outcome, iterations) to FILE\n"
              << "  -h, --help       this help\n"
              << "  -V, --version    version\n";
}

}  // namespace

int main(int argc, char *argv[])
{
    bool verbose = false, fixed_seed = false, check_only = false;
    unsigned seed = 0;
    std::string file, steps_log;
    for (int k = 1; k < argc; k++) {
        const std::string a = argv[k];
        if (a == "-v" || a == "--verbose") verbose = true;
        else if (a == "--check") check_only = true;
        else if (a == "--seed" && k + 1 < argc) {
            fixed_seed = true;
            seed = (unsigned)std::strtoul(argv[++k], nullptr, 10);
        } else if (a == "--steps-log" && k + 1 < argc) {
            steps_log = argv[++k];
        } else if (a == "-h" || a == "--help") {
            usage(argv[0]);
            return 0;
        } else if (a == "-V" || a == "--version") {
            std::cout << VERSION << "\n";
            return 0;
        } else if (a == "-" || a[0] != '-') {
            if (!file.empty()) {
                usage(argv[0]);
                return 1;
            }
            file = a;
        } else {
            std::cerr << "unknown option " << a << "\n";
            usage(argv[0]);
            return 1;
        }
    }
    if (file.empty()) {
        usage(argv[0]);
        return 1;
    }
    if (!fixed_seed) seed = std::random_device{}();
    try {
        std::string text;
        if (file == "-") {
            std::ostringstream ss;
            ss << std::cin.rdbuf();
            text = ss.str();
        } else {
            std::ifstream in(file);
            if (!in) throw std::runtime_error("cannot open " + file);
            std::ostringstream ss;
            ss << in.rdbuf();
            text = ss.str();
        }
        fgh::Node root = fgh::parse_settings(text);
        {  // the user's defaults first, then the settings over them
            const std::string dname = user_defaults_filename();
            std::ifstream din(dname);
            if (din) {
                std::ostringstream ds;
                ds << din.rdbuf();
                fgh::Node defaults;
                try {
                    defaults = fgh::parse_settings(ds.str());
                } catch (const std::exception &ex) {
                    throw std::runtime_error(dname + ": " + ex.what());
                }
                if (!defaults.is_null()) root = merge_settings(defaults, root);
            }
        }
        Config c = read_config(root);
        c.settings_name = file == "-" ? "standard input" : file;
        std::cout << "random seed:       " << seed << "\n";
        // SIGINT and SIGTERM save the state before terminating (src/main.cpp:256-271)
        struct sigaction action;
        action.sa_handler = signal_handler;
        action.sa_flags = 0;
        sigemptyset(&action.sa_mask);
        if (sigaction(SIGINT, &action, nullptr) == -1 || sigaction(SIGTERM, &action, nullptr) == -1) {
            perror("sigaction");
            return 1;
        }
        return run(c, seed, verbose, check_only, steps_log);
    } catch (const std::exception &ex) {
        std::cerr << "error: " << ex.what() << "\n";
        return 1;
    }
}
