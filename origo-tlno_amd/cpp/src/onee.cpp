// One-electron API of the C++ facade (tlno::integrals::onee): the
// reference's host plan (shell groups, screened group pairs, primitive-pair
// data, spherical layout; integrals/one_electron/onee_matrices.cpp:47-216,
// onee_integrals.cpp:160-259) built on the host as the reference builds it,
// and every S/T/V evaluation done by the GPU engine (tlno_gpu_onee_screened)
// under the same screening rules.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/basis/types.hpp"
#include "facade_internal.hpp"
#include "integrals/one_electron/onee_integrals.hpp"
#include "integrals/one_electron/onee_matrices.hpp"
#include "tlno_gpu.h"

namespace tlno::integrals::onee {

using tlno_amd_facade::check;
using tlno_amd_facade::GpuPlan;
using tlno_amd_facade::make_gpu_plan;

namespace {

constexpr double kGroupPairDecayEps = 1.0e-15;
double g_decay_eps = kGroupPairDecayEps;
double g_nuclear_eps = kNuclearScreenEps;

int ncart(int L) { return (L + 1) * (L + 2) / 2; }

double dist2(const std::array<double, 3>& a, const std::array<double, 3>& b) {
    const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return x * x + y * y + z * z;
}

// Components of one centre, angular momentum and exponent list form a group
// (onee_matrices.cpp:47-95): first-appearance order of groups and of the
// components inside a group.
std::vector<ShellGroup> make_groups(const basis::BasisBuildResult& build) {
    std::vector<ShellGroup> groups;
    for (int i = 0; i < static_cast<int>(build.basis.size()); ++i) {
        const basis::Shell& sh = build.basis[i];
        const int L = sh.ang[0] + sh.ang[1] + sh.ang[2];
        if (sh.primitives.empty()) throw std::runtime_error("onee: contracted shells need at least one primitive.");
        std::vector<double> ex, cf;
        for (const auto& p : sh.primitives) {
            ex.push_back(p.exponent);
            cf.push_back(p.coefficient);
        }
        auto it = std::find_if(groups.begin(), groups.end(), [&](const ShellGroup& g) {
            return g.L == L && g.center == sh.center && g.exponents == ex;
        });
        if (it == groups.end()) {
            ShellGroup g;
            g.center = sh.center;
            g.L = L;
            g.exponents = ex;
            groups.push_back(std::move(g));
            it = groups.end() - 1;
        }
        it->basis_indices.push_back(i);
        it->ang.push_back(sh.ang);
        it->coeffs.push_back(std::move(cf));
    }
    return groups;
}

// Group pairs and their flags (onee_matrices.cpp:97-210): with c0 the first
// component's coefficients, K_g = sum_i |c0_i| (pi/a_i)^{3/4}; a pair is
// dropped when K_a K_b exp(-mu_min AB^2) < eps, and evaluates V only when
// its overlap bound sum_ij |c0a_i c0b_j| (pi/p)^{3/2} exp(-chi AB^2) reaches
// eps; the higher-L group goes left.
std::vector<OneElectronGroupPair> make_group_pairs(const std::vector<ShellGroup>& groups, double eps) {
    const int n = static_cast<int>(groups.size());
    std::vector<double> K(n, 0.0), amin(n, 0.0);
    for (int g = 0; g < n; ++g) {
        const ShellGroup& G = groups[g];
        amin[g] = *std::min_element(G.exponents.begin(), G.exponents.end());
        for (std::size_t i = 0; i < G.exponents.size(); ++i)
            K[g] += std::abs(G.coeffs[0][i]) * std::pow(basis::kPi / G.exponents[i], 0.75);
    }
    std::vector<OneElectronGroupPair> out;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b <= a; ++b) {
            OneElectronGroupPair gp;
            gp.ga = a;
            gp.gb = b;
            if (eps > 0.0) {
                const double AB2 = dist2(groups[a].center, groups[b].center);
                const double mu = amin[a] * amin[b] / (amin[a] + amin[b]);
                if (K[a] * K[b] * std::exp(-mu * AB2) < eps) continue;
                double bound = 0.0;
                const ShellGroup &A = groups[a], &B = groups[b];
                for (std::size_t i = 0; i < A.exponents.size(); ++i)
                    for (std::size_t j = 0; j < B.exponents.size(); ++j) {
                        const double p = A.exponents[i] + B.exponents[j];
                        const double chi = A.exponents[i] * B.exponents[j] / p;
                        bound += std::abs(A.coeffs[0][i] * B.coeffs[0][j]) * std::pow(basis::kPi / p, 1.5) *
                                 std::exp(-chi * AB2);
                    }
                gp.do_v = bound >= eps;
            }
            if (groups[gp.gb].L > groups[gp.ga].L) std::swap(gp.ga, gp.gb);
            out.push_back(gp);
        }
    return out;
}

// Spherical AO base of every Cartesian shell (chunk of ncart(L) components)
// of every group, assigned in the order of the shells' first Cartesian
// component (onee_matrices.cpp:459-484).
void spherical_layout(OneElectronPlan& plan) {
    struct Chunk {
        int start, g, c, L;
    };
    std::vector<Chunk> chunks;
    plan.sph_base.assign(plan.groups.size(), {});
    for (int g = 0; g < static_cast<int>(plan.groups.size()); ++g) {
        const ShellGroup& G = plan.groups[g];
        const int nc = ncart(G.L);
        const int nchunk = static_cast<int>(G.basis_indices.size()) / nc;
        plan.sph_base[g].assign(nchunk, 0);
        for (int c = 0; c < nchunk; ++c) chunks.push_back({G.basis_indices[c * nc], g, c, G.L});
    }
    std::sort(chunks.begin(), chunks.end(), [](const Chunk& x, const Chunk& y) { return x.start < y.start; });
    int base = 0;
    for (const Chunk& ch : chunks) {
        plan.sph_base[ch.g][ch.c] = base;
        base += 2 * ch.L + 1;
    }
    plan.n_sph = base;
}

std::vector<double> charges_of(const basis::BasisBuildResult& build, std::vector<double>& R) {
    if (build.nuclear_charges.size() != build.nuclear_centers.size())
        throw std::runtime_error("onee: nuclear charges and centres differ in length.");
    R.clear();
    for (const auto& c : build.nuclear_centers) R.insert(R.end(), c.begin(), c.end());
    return std::vector<double>(build.nuclear_charges.begin(), build.nuclear_charges.end());
}

}  // namespace

// GPU plans of the basis (Cartesian and spherical, created on first use) and
// the decay threshold in force when the plan was built.
struct OneElectronPlanBackend {
    std::vector<basis::Shell> shells;
    double decay_eps = 0.0;
    mutable std::mutex mu;
    mutable std::shared_ptr<GpuPlan> cart, sph;
    GpuPlan& get(bool spherical) const {
        std::lock_guard<std::mutex> lk(mu);
        auto& p = spherical ? sph : cart;
        if (!p) p = make_gpu_plan(shells, spherical ? 1 : 0);
        return *p;
    }
};

ShellPairPrim build_shell_pair_prim(const ShellGroup& ga, const ShellGroup& gb) {
    ShellPairPrim spp;
    const int na = static_cast<int>(ga.exponents.size()), nb = static_cast<int>(gb.exponents.size());
    spp.n_prim_a = na;
    spp.n_prim_b = nb;
    spp.ABx = ga.center[0] - gb.center[0];
    spp.ABy = ga.center[1] - gb.center[1];
    spp.ABz = ga.center[2] - gb.center[2];
    spp.AB2 = spp.ABx * spp.ABx + spp.ABy * spp.ABy + spp.ABz * spp.ABz;
    const std::size_t npp = static_cast<std::size_t>(na) * nb;
    const std::size_t sec = (npp + 7) & ~static_cast<std::size_t>(7);
    spp.storage.assign(18 * sec, 0.0);
    double** fields[18] = {&spp.p,   &spp.inv2p, &spp.chi, &spp.sss, &spp.expmChiAB2, &spp.alpha,
                           &spp.beta, &spp.PAx,  &spp.PAy, &spp.PAz, &spp.Px,         &spp.Py,
                           &spp.Pz,  &spp.e0x,   &spp.e0y, &spp.e0z, &spp.neg2pi_over_p, &spp.screen};
    for (int f = 0; f < 18; ++f) *fields[f] = spp.storage.data() + f * sec;
    std::vector<double> mca(na, 0.0), mcb(nb, 0.0);
    for (const auto& row : ga.coeffs)
        for (int i = 0; i < na; ++i) mca[i] = std::max(mca[i], std::abs(row[i]));
    for (const auto& row : gb.coeffs)
        for (int j = 0; j < nb; ++j) mcb[j] = std::max(mcb[j], std::abs(row[j]));
    const double AB[3] = {spp.ABx, spp.ABy, spp.ABz};
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            const std::size_t k = static_cast<std::size_t>(i) * nb + j;
            const double a = ga.exponents[i], b = gb.exponents[j], p = a + b, ip = 1.0 / p;
            const double chi = a * b * ip, emc = std::exp(-chi * spp.AB2), pop = basis::kPi * ip;
            spp.p[k] = p;
            spp.inv2p[k] = 0.5 * ip;
            spp.chi[k] = chi;
            spp.alpha[k] = a;
            spp.beta[k] = b;
            spp.expmChiAB2[k] = emc;
            spp.sss[k] = pop * std::sqrt(pop) * emc;
            double* P[3] = {spp.Px, spp.Py, spp.Pz};
            double* PA[3] = {spp.PAx, spp.PAy, spp.PAz};
            double* e0[3] = {spp.e0x, spp.e0y, spp.e0z};
            for (int x = 0; x < 3; ++x) {
                const double px = (a * ga.center[x] + b * gb.center[x]) * ip;
                P[x][k] = px;
                PA[x][k] = px - ga.center[x];
                e0[x][k] = std::exp(-chi * AB[x] * AB[x]);
            }
            spp.neg2pi_over_p[k] = -2.0 * pop;
            spp.screen[k] = 2.0 * pop * emc * mca[i] * mcb[j];
        }
    return spp;
}

// The GPU engine evaluates V by Rys quadrature from the pair data; there is
// no Hermite projector to cache, and an empty cache is the reference's
// "not cached" state.
void cache_shell_pair_nucleus_hermite(const ShellGroup&, const ShellGroup&, ShellPairPrim& spp) {
    spp.nuclear_hermite.clear();
}

namespace {

// S, T and/or V of the block (ga, gb) on the GPU: a Cartesian two-group plan
// of the groups' components, evaluated whole; the [ga][gb] block is added to
// the caller's blocks.
void eval_block(const ShellGroup& ga, const ShellGroup& gb, double* S, double* T, double* V,
                const std::vector<int>* charges, const std::vector<std::array<double, 3>>* centers) {
    const bool same = &ga == &gb || (ga.L == gb.L && ga.center == gb.center && ga.exponents == gb.exponents &&
                                     ga.ang == gb.ang && ga.coeffs == gb.coeffs);
    std::vector<basis::Shell> shells;
    auto add = [&](const ShellGroup& g) {
        if (g.ang.size() != g.coeffs.size()) throw std::runtime_error("onee: shell group components and coefficients differ in length.");
        for (std::size_t c = 0; c < g.ang.size(); ++c) {
            basis::Shell s;
            s.center = g.center;
            s.ang = g.ang[c];
            if (g.coeffs[c].size() != g.exponents.size())
                throw std::runtime_error("onee: shell group exponents and coefficients differ in length.");
            for (std::size_t i = 0; i < g.exponents.size(); ++i) s.primitives.push_back({g.exponents[i], g.coeffs[c][i]});
            shells.push_back(std::move(s));
        }
    };
    add(ga);
    if (!same) add(gb);
    const int na = static_cast<int>(ga.ang.size()), nb = static_cast<int>(gb.ang.size());
    const int b0 = same ? 0 : na;
    auto plan = make_gpu_plan(shells, 0);
    const int n = plan->M;
    std::vector<double> Z, R;
    if (charges) {
        if (charges->size() != centers->size())
            throw std::runtime_error("onee: nuclear charges and centres differ in length.");
        Z.assign(charges->begin(), charges->end());
        for (const auto& c : *centers) R.insert(R.end(), c.begin(), c.end());
    }
    std::vector<double> mS(S ? static_cast<std::size_t>(n) * n : 0), mT(T ? mS.size() : 0),
        mV(V ? static_cast<std::size_t>(n) * n : 0);
    check(tlno_gpu_onee_screened(plan->h, static_cast<int>(Z.size()), Z.data(), R.data(), 0.0, g_nuclear_eps,
                                 S ? mS.data() : nullptr, T ? mT.data() : nullptr, V ? mV.data() : nullptr,
                                 nullptr));
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            const std::size_t o = static_cast<std::size_t>(i) * n + b0 + j;
            if (S) S[i * nb + j] += mS[o];
            if (T) T[i * nb + j] += mT[o];
            if (V) V[i * nb + j] += mV[o];
        }
}

}  // namespace

void contracted_shell_block_overlap_kinetic(const ShellGroup& ga, const ShellGroup& gb, const ShellPairPrim&,
                                            double* overlap_block, double* kinetic_block, double min_abs_overlap) {
    if (min_abs_overlap != 0.0)
        throw std::runtime_error("onee: the GPU engine evaluates every primitive pair (min_abs_overlap must be 0).");
    eval_block(ga, gb, overlap_block, kinetic_block, nullptr, nullptr, nullptr);
}

void contracted_shell_block_nuclear(const ShellGroup& ga, const ShellGroup& gb, const ShellPairPrim&,
                                    double* nuclear_block, const std::vector<int>& nuclear_charges,
                                    const std::vector<std::array<double, 3>>& nuclear_centers) {
    eval_block(ga, gb, nullptr, nullptr, nuclear_block, &nuclear_charges, &nuclear_centers);
}

void set_onee_nuclear_screen_eps(double eps) { g_nuclear_eps = eps; }
double get_onee_nuclear_screen_eps() { return g_nuclear_eps; }

namespace {

OneElectronPlan make_plan(const basis::BasisBuildResult& build, bool spherical, bool with_pair_prims) {
    OneElectronPlan plan;
    plan.groups = make_groups(build);
    plan.group_pairs = make_group_pairs(plan.groups, g_decay_eps);
    if (with_pair_prims) {
        plan.pair_prims.reserve(plan.group_pairs.size());
        for (const auto& gp : plan.group_pairs)
            plan.pair_prims.push_back(build_shell_pair_prim(plan.groups[gp.ga], plan.groups[gp.gb]));
    }
    spherical_layout(plan);
    auto b = std::make_shared<OneElectronPlanBackend>();
    b->shells = build.basis;
    b->decay_eps = g_decay_eps;
    (void)b->get(spherical);
    plan.backend = std::move(b);
    return plan;
}

void evaluate(const basis::BasisBuildResult& build, const OneElectronPlan& plan, bool spherical,
              OneElectronMatrices& mats) {
    if (!plan.backend) throw std::runtime_error("onee: null one-electron plan.");
    GpuPlan& g = plan.backend->get(spherical);
    std::vector<double> R;
    const std::vector<double> Z = charges_of(build, R);
    if (mats.overlap.size != g.M) mats = OneElectronMatrices(g.M);
    std::lock_guard<std::mutex> lk(g.mu);
    check(tlno_gpu_onee_screened(g.h, static_cast<int>(Z.size()), Z.data(), R.data(), plan.backend->decay_eps,
                                 g_nuclear_eps, mats.overlap.values.data(), mats.kinetic.values.data(),
                                 mats.nuclear.values.data(), nullptr));
}

}  // namespace

OneElectronPlan build_one_electron_plan(const basis::BasisBuildResult& build) {
    return make_plan(build, false, true);
}
OneElectronPlan build_one_electron_plan(const basis::BasisBuildResult& build, bool) {
    return make_plan(build, false, true);
}
OneElectronPlan build_one_electron_spherical_plan(const basis::BasisBuildResult& build, bool) {
    return make_plan(build, true, true);
}

OneElectronMatrices evaluate_one_electron_matrices(const basis::BasisBuildResult& build, const OneElectronPlan& plan) {
    OneElectronMatrices m;
    evaluate(build, plan, false, m);
    return m;
}

void evaluate_one_electron_matrices_into(const basis::BasisBuildResult& build, const OneElectronPlan& plan,
                                         OneElectronMatrices& mats) {
    evaluate(build, plan, false, mats);
}

void evaluate_one_electron_matrices_sph_into(const basis::BasisBuildResult& build, const OneElectronPlan& plan,
                                             OneElectronMatrices& mats) {
    evaluate(build, plan, true, mats);
}

OneElectronMatrices build_one_electron_matrices(const basis::BasisBuildResult& build) {
    return evaluate_one_electron_matrices(build, make_plan(build, false, false));
}

OneElectronMatrices build_one_electron_spherical_matrices(const basis::BasisBuildResult& build) {
    OneElectronMatrices m;
    evaluate(build, make_plan(build, true, false), true, m);
    return m;
}

void set_onee_screening_enabled(bool enabled) {
    g_decay_eps = enabled ? kGroupPairDecayEps : 0.0;
    g_nuclear_eps = enabled ? kNuclearScreenEps : 0.0;
}
bool onee_screening_enabled() { return g_decay_eps > 0.0; }
void set_onee_group_pair_decay_eps(double eps) { g_decay_eps = eps; }
double onee_group_pair_decay_eps() { return g_decay_eps; }

}  // namespace tlno::integrals::onee
