// The reference's one-electron C ABI (tlno_onee_*,
// benchmarks/one_electron_engine/src/onee_capi.cpp:36-225 and
// basis_export.cpp:43) over the facade's one-electron API, so the
// reference's ctypes harness loads this library unchanged.  Declared in
// include/tlno_onee.h.
//
// A handle owns the molecule's basis, the reusable plan and the output
// matrices; the eval calls write into the handle and copy hands them out.
// Basis input: the named basis tables of the reference are not part of this
// build, so basis_name names a JSON basis file ("*.json": per element symbol
// {"charge": Z, "shells": [[L, [exponent, coefficient], ...], ...]}, exactly
// the objects tlno_onee_basis_json returns); each radial shell expands to its
// Cartesian components in descending-lex order with per-component
// normalization (core/shell/shell_contract.cpp:9-48).  Named bases return
// nullptr, as the reference does for unknown names.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/basis/types.hpp"
#include "integrals/one_electron/onee_matrices.hpp"
#include "tlno_onee.h"

namespace {

namespace basis = tlno::core::basis;
namespace onee = tlno::integrals::onee;

// ---- minimal JSON reader (objects, arrays, numbers, strings) ----
struct Json {
    enum Kind { Null, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
};

struct JsonParser {
    const std::string& t;
    std::size_t i = 0;
    explicit JsonParser(const std::string& text) : t(text) {}
    void ws() {
        while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) ++i;
    }
    [[noreturn]] void fail() { throw std::runtime_error("basis JSON: syntax error at offset " + std::to_string(i)); }
    std::string string() {
        if (t[i] != '"') fail();
        std::string s;
        for (++i; i < t.size() && t[i] != '"'; ++i) {
            if (t[i] == '\\' && i + 1 < t.size()) ++i;
            s += t[i];
        }
        if (i >= t.size()) fail();
        ++i;
        return s;
    }
    Json value() {
        ws();
        if (i >= t.size()) fail();
        Json v;
        if (t[i] == '{') {
            v.kind = Json::Obj;
            ++i;
            ws();
            if (t[i] == '}') {
                ++i;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (t[i++] != ':') fail();
                v.obj.emplace_back(k, value());
                ws();
                if (t[i] == ',') {
                    ++i;
                    continue;
                }
                if (t[i++] != '}') fail();
                return v;
            }
        }
        if (t[i] == '[') {
            v.kind = Json::Arr;
            ++i;
            ws();
            if (t[i] == ']') {
                ++i;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (t[i] == ',') {
                    ++i;
                    continue;
                }
                if (t[i++] != ']') fail();
                return v;
            }
        }
        if (t[i] == '"') {
            v.kind = Json::Str;
            v.str = string();
            return v;
        }
        const char* b = t.c_str() + i;
        char* e = nullptr;
        v.num = std::strtod(b, &e);
        if (e == b) fail();
        v.kind = Json::Num;
        i += static_cast<std::size_t>(e - b);
        return v;
    }
};

struct RadialShell {
    int L;
    std::vector<double> exps, coefs;
};
struct ElementBasis {
    int charge = 0;
    std::vector<RadialShell> shells;
};

bool is_json_path(const std::string& name) {
    return name.size() > 5 && name.compare(name.size() - 5, 5, ".json") == 0;
}

std::string upper(std::string s) {
    for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

std::map<std::string, ElementBasis> read_basis_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open basis file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JsonParser jp(text);
    const Json root = jp.value();
    if (root.kind != Json::Obj) throw std::runtime_error("basis JSON: top level must be an object");
    std::map<std::string, ElementBasis> out;
    for (const auto& [sym, el] : root.obj) {
        ElementBasis eb;
        for (const auto& [k, v] : el.obj) {
            if (k == "charge") eb.charge = static_cast<int>(v.num);
            if (k != "shells") continue;
            for (const Json& sh : v.arr) {
                if (sh.kind != Json::Arr || sh.arr.size() < 2) throw std::runtime_error("basis JSON: bad shell");
                RadialShell r;
                r.L = static_cast<int>(sh.arr[0].num);
                for (std::size_t p = 1; p < sh.arr.size(); ++p) {
                    if (sh.arr[p].arr.size() != 2) throw std::runtime_error("basis JSON: bad primitive");
                    r.exps.push_back(sh.arr[p].arr[0].num);
                    r.coefs.push_back(sh.arr[p].arr[1].num);
                }
                eb.shells.push_back(std::move(r));
            }
        }
        out[upper(sym)] = std::move(eb);
    }
    return out;
}

// ".inp" molecule (benchmarks/common/molecule_io.cpp:120-165): lines before a
// "* xyz charge multiplicity" header are ignored, then "Symbol x y z" in bohr
// ('#' or '!' comment lines skipped) until a line starting with '*'.
struct Molecule {
    int charge = 0;
    std::vector<std::string> symbols;
    std::vector<std::array<double, 3>> xyz;
};

std::string trim(const std::string& s) {
    std::size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

Molecule read_inp(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Cannot open input file: " + path);
    Molecule m;
    bool open = false, closed = false;
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        const std::string t = trim(line);
        if (t.empty()) continue;
        if (!open) {
            if (t[0] != '*') continue;
            std::istringstream h(t);
            char star = 0;
            std::string kw;
            int q = 0, mult = 0;
            if ((h >> star >> kw >> q >> mult) && star == '*' && (kw == "xyz" || kw == "XYZ" || kw == "Xyz")) {
                if (mult < 1) throw std::runtime_error("xyz multiplicity must be at least one: " + path);
                m.charge = q;
                open = true;
            }
            continue;
        }
        if (t[0] == '*') {
            closed = true;
            break;
        }
        if (t[0] == '#' || t[0] == '!') continue;
        std::istringstream a(t);
        std::string sym;
        std::array<double, 3> r{};
        if (!(a >> sym >> r[0] >> r[1] >> r[2]) || !std::isfinite(r[0]) || !std::isfinite(r[1]) ||
            !std::isfinite(r[2]))
            throw std::runtime_error("bad coordinate line " + std::to_string(ln) + ": " + path);
        m.symbols.push_back(sym);
        m.xyz.push_back(r);
    }
    if (!open) throw std::runtime_error("Missing '* xyz charge multiplicity' block in input: " + path);
    if (!closed) throw std::runtime_error("Unterminated xyz block: " + path);
    if (m.symbols.empty()) throw std::runtime_error("Empty xyz block: " + path);
    return m;
}

double dfact(int n) {  // n!! for n >= -1
    double v = 1.0;
    for (int k = n; k > 1; k -= 2) v *= k;
    return v;
}

// Normalized primitives of one Cartesian component: each primitive gets
// N = sqrt(2^{2L+3/2} a^{L+3/2} / (pi^{3/2} (2lx-1)!!(2ly-1)!!(2lz-1)!!)),
// then the contraction is scaled to unit self-overlap with the same-centre
// metric (pi/p)^{3/2} prod_x (2l_x-1)!!/(2p)^{l_x}.
std::vector<basis::Primitive> normalized(const RadialShell& r, const std::array<int, 3>& ang) {
    const int L = ang[0] + ang[1] + ang[2];
    const double df = dfact(2 * ang[0] - 1) * dfact(2 * ang[1] - 1) * dfact(2 * ang[2] - 1);
    std::vector<basis::Primitive> p;
    for (std::size_t i = 0; i < r.exps.size(); ++i) {
        const double a = r.exps[i];
        const double N = std::sqrt(std::pow(2.0, 2.0 * L + 1.5) * std::pow(a, L + 1.5) / (std::pow(basis::kPi, 1.5) * df));
        p.push_back({a, r.coefs[i] * N});
    }
    double norm2 = 0.0;
    for (const auto& x : p)
        for (const auto& y : p) {
            const double s = x.exponent + y.exponent;
            double m = std::pow(basis::kPi / s, 1.5);
            for (int k = 0; k < 3; ++k) m *= dfact(2 * ang[k] - 1) / std::pow(2.0 * s, ang[k]);
            norm2 += x.coefficient * y.coefficient * m;
        }
    const double sc = 1.0 / std::sqrt(norm2);
    for (auto& x : p) x.coefficient *= sc;
    return p;
}

basis::BasisBuildResult build_from(const Molecule& mol, const std::map<std::string, ElementBasis>& lib) {
    basis::BasisBuildResult br;
    br.nuclear_centers = mol.xyz;
    for (std::size_t a = 0; a < mol.symbols.size(); ++a) {
        auto it = lib.find(upper(mol.symbols[a]));
        if (it == lib.end()) throw std::runtime_error("basis file has no element " + mol.symbols[a]);
        const ElementBasis& eb = it->second;
        br.nuclear_charges.push_back(eb.charge);
        br.electron_count += eb.charge;
        for (const RadialShell& r : eb.shells)
            for (int x = r.L; x >= 0; --x)
                for (int y = r.L - x; y >= 0; --y) {
                    basis::Shell s;
                    s.center = mol.xyz[a];
                    s.ang = {x, y, r.L - x - y};
                    s.primitives = normalized(r, s.ang);
                    br.basis.push_back(std::move(s));
                }
    }
    br.electron_count -= mol.charge;
    return br;
}

struct Handle {
    basis::BasisBuildResult br;
    onee::OneElectronPlan plan;
    onee::OneElectronMatrices out, out_sph;
    bool has_plan = false;
};

Handle* H(void* h) { return static_cast<Handle*>(h); }

}  // namespace

extern "C" {

void* tlno_onee_create_from_inp(const char* inp_path, const char* basis_name) {
    try {
        if (!inp_path || !basis_name || !is_json_path(basis_name)) return nullptr;
        auto h = std::make_unique<Handle>();
        h->br = build_from(read_inp(inp_path), read_basis_file(basis_name));
        h->out = onee::OneElectronMatrices(static_cast<int>(h->br.basis.size()));
        return h.release();
    } catch (...) {
        return nullptr;
    }
}

void* tlno_onee_create(int n_comp, const double* centers, const int* ang, const int* nprim, const double* exps,
                       const double* coefs, int n_nuc, const int* charges, const double* nuc_centers) {
    try {
        if (n_comp <= 0 || !centers || !ang || !nprim || !exps || !coefs || n_nuc < 0 ||
            (n_nuc > 0 && (!charges || !nuc_centers)))
            return nullptr;
        auto h = std::make_unique<Handle>();
        std::size_t off = 0;
        for (int i = 0; i < n_comp; ++i) {
            basis::Shell s;
            for (int x = 0; x < 3; ++x) {
                s.center[x] = centers[3 * i + x];
                s.ang[x] = ang[3 * i + x];
            }
            for (int k = 0; k < nprim[i]; ++k, ++off) s.primitives.push_back({exps[off], coefs[off]});
            h->br.basis.push_back(std::move(s));
        }
        for (int c = 0; c < n_nuc; ++c) {
            h->br.nuclear_charges.push_back(charges[c]);
            h->br.nuclear_centers.push_back({nuc_centers[3 * c], nuc_centers[3 * c + 1], nuc_centers[3 * c + 2]});
            h->br.electron_count += charges[c];
        }
        h->out = onee::OneElectronMatrices(n_comp);
        return h.release();
    } catch (...) {
        return nullptr;
    }
}

int tlno_onee_nbf(void* handle) { return handle ? static_cast<int>(H(handle)->br.basis.size()) : -1; }

int tlno_onee_nbf_sph(void* handle) {
    if (!handle) return -1;
    try {
        Handle* h = H(handle);
        if (!h->has_plan) {
            h->plan = onee::build_one_electron_spherical_plan(h->br);
            h->has_plan = true;
        }
        return h->plan.n_sph;
    } catch (...) {
        return -1;
    }
}

void tlno_onee_set_screening(int enabled) { onee::set_onee_screening_enabled(enabled != 0); }
int tlno_onee_screening_enabled(void) { return onee::onee_screening_enabled() ? 1 : 0; }
void tlno_onee_set_group_pair_decay_eps(double eps) { onee::set_onee_group_pair_decay_eps(eps); }
double tlno_onee_group_pair_decay_eps(void) { return onee::onee_group_pair_decay_eps(); }
void tlno_onee_set_nuclear_screen_eps(double eps) { onee::set_onee_nuclear_screen_eps(eps); }
double tlno_onee_nuclear_screen_eps(void) { return onee::get_onee_nuclear_screen_eps(); }

void tlno_onee_prepare_plan(void* handle) {
    if (!handle) return;
    try {
        H(handle)->plan = onee::build_one_electron_plan(H(handle)->br, true);
        H(handle)->has_plan = true;
    } catch (...) {
    }
}

void tlno_onee_eval_full(void* handle) {
    if (!handle) return;
    try {
        Handle* h = H(handle);
        h->plan = onee::build_one_electron_plan(h->br, false);
        h->has_plan = true;
        onee::evaluate_one_electron_matrices_into(h->br, h->plan, h->out);
    } catch (...) {
    }
}

void tlno_onee_eval_cached(void* handle) {
    if (!handle || !H(handle)->has_plan) return;
    try {
        onee::evaluate_one_electron_matrices_into(H(handle)->br, H(handle)->plan, H(handle)->out);
    } catch (...) {
    }
}

void tlno_onee_eval_full_sph(void* handle) {
    if (!handle) return;
    try {
        Handle* h = H(handle);
        h->plan = onee::build_one_electron_spherical_plan(h->br, false);
        h->has_plan = true;
        onee::evaluate_one_electron_matrices_sph_into(h->br, h->plan, h->out_sph);
    } catch (...) {
    }
}

void tlno_onee_eval_cached_sph(void* handle) {
    if (!handle || !H(handle)->has_plan) return;
    try {
        onee::evaluate_one_electron_matrices_sph_into(H(handle)->br, H(handle)->plan, H(handle)->out_sph);
    } catch (...) {
    }
}

static void copy3(const onee::OneElectronMatrices& m, double* S, double* T, double* V) {
    const std::size_t n = m.overlap.values.size() * sizeof(double);
    if (S && n) std::memcpy(S, m.overlap.values.data(), n);
    if (T && n) std::memcpy(T, m.kinetic.values.data(), n);
    if (V && n) std::memcpy(V, m.nuclear.values.data(), n);
}

void tlno_onee_copy(void* handle, double* S, double* T, double* V) {
    if (handle) copy3(H(handle)->out, S, T, V);
}

void tlno_onee_copy_sph(void* handle, double* S, double* T, double* V) {
    if (handle) copy3(H(handle)->out_sph, S, T, V);
}

void tlno_onee_pair_stats(void* handle, int* n_groups, int* total_possible, int* kept) {
    if (!handle) return;
    const onee::OneElectronPlan& p = H(handle)->plan;
    const int g = static_cast<int>(p.groups.size());
    if (n_groups) *n_groups = g;
    if (total_possible) *total_possible = g * (g + 1) / 2;
    if (kept) *kept = static_cast<int>(p.group_pairs.size());
}

void tlno_onee_nuclear_pair_stats(void* handle, long long* total, long long* kept) {
    if (!handle) return;
    const Handle* h = H(handle);
    double zsum = 0.0;
    for (int z : h->br.nuclear_charges) zsum += std::abs(static_cast<double>(z));
    const double thr = zsum > 0.0 ? onee::get_onee_nuclear_screen_eps() / zsum : 0.0;
    long long t = 0, k = 0;
    for (std::size_t i = 0; i < h->plan.group_pairs.size() && i < h->plan.pair_prims.size(); ++i) {
        if (!h->plan.group_pairs[i].do_v) continue;
        const onee::ShellPairPrim& spp = h->plan.pair_prims[i];
        const long long n = static_cast<long long>(spp.n_prim_a) * spp.n_prim_b;
        t += n;
        for (long long q = 0; q < n; ++q)
            if (spp.screen[q] >= thr) ++k;
    }
    if (total) *total = t;
    if (kept) *kept = k;
}

void tlno_onee_noop(void* handle) { (void)handle; }

void tlno_onee_destroy(void* handle) { delete H(handle); }

const char* tlno_onee_basis_json(const char* basis_name, const char* symbol) {
    thread_local std::string result;
    try {
        if (!basis_name || !symbol || !is_json_path(basis_name)) return nullptr;
        const auto lib = read_basis_file(basis_name);
        auto it = lib.find(upper(symbol));
        if (it == lib.end()) return nullptr;
        std::ostringstream js;
        js.precision(17);
        js << "{\"charge\":" << it->second.charge << ",\"shells\":[";
        for (std::size_t s = 0; s < it->second.shells.size(); ++s) {
            const RadialShell& r = it->second.shells[s];
            js << (s ? "," : "") << '[' << r.L;
            for (std::size_t p = 0; p < r.exps.size(); ++p) js << ",[" << r.exps[p] << ',' << r.coefs[p] << ']';
            js << ']';
        }
        js << "]}";
        result = js.str();
        return result.c_str();
    } catch (...) {
        result.clear();
        return nullptr;
    }
}

}  // extern "C"
