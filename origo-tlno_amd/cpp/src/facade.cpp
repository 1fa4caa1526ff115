// C++ facade over the C ABI (include/tlno_gpu.h): the reference's one- and
// two-electron engine API (namespaces tlno::core, tlno::integrals::twoe,
// tlno::integrals::twoe::fock, tlno::integrals::onee) implemented on the GPU
// engine.  Errors raise std::runtime_error with the engine's message text.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/basis/types.hpp"
#include "core/math/general_math.hpp"
#include "integrals/two_electron/fock/eri_direct_jk.hpp"
#include "integrals/two_electron/fock/jk_direct.hpp"
#include "integrals/two_electron/twoe_integrals.hpp"
#include "facade_internal.hpp"
#include "tlno_gpu.h"

namespace {

using tlno::core::basis::Shell;
using tlno_amd_facade::check;
using tlno_amd_facade::finite_nonneg;
using tlno_amd_facade::GpuPlan;
using tlno_amd_facade::make_gpu_plan;

void check_heuristic_screens(double fastschwarz, double decay) {
    if (!finite_nonneg(fastschwarz))
        throw std::runtime_error("eri_ao_tensor_hgp: primitive_fastschwarz_threshold must be finite and >= 0.");
    if (!finite_nonneg(decay)) throw std::runtime_error("eri_ao_tensor_hgp: decay_threshold must be finite and >= 0.");
}

// The stream checksum takes the Schwarz threshold only.
void require_schwarz_only(double fastschwarz, double decay) {
    check_heuristic_screens(fastschwarz, decay);
    if (fastschwarz != 0.0 || decay != 0.0)
        throw std::runtime_error(
            "eri_ao_tensor_hgp: the GPU stream checksum applies Schwarz screening only (primitive FASTSCHWARZ and "
            "decay thresholds must be 0).");
}

// All screening thresholds of one J/K build.
struct Screens {
    double schwarz, density, fastschwarz, decay;
};

// J and/or K on a locked plan
void build_jk(GpuPlan& g, const std::vector<double>& density, const Screens& sc, double* J, double* K) {
    if (density.size() != static_cast<std::size_t>(g.M) * g.M)
        throw std::runtime_error("eri_ao_tensor_hgp: density must be spherical M x M row-major.");
    if (!finite_nonneg(sc.density))
        throw std::runtime_error("eri_ao_tensor_hgp: density_threshold must be finite and >= 0.");
    check_heuristic_screens(sc.fastschwarz, sc.decay);
    std::lock_guard<std::mutex> lk(g.mu);
    check(tlno_gpu_set_density_threshold(g.h, sc.density));
    check(tlno_gpu_set_primitive_screens(g.h, sc.fastschwarz, sc.decay));
    check(tlno_gpu_build_jk(g.h, density.data(), J, K, sc.schwarz));
}

tlno::integrals::twoe::EriPackedPlanProfile profile_of(const GpuPlan& g) {
    tlno::integrals::twoe::EriPackedPlanProfile p;
    p.M = g.M;
    p.shell_groups = static_cast<std::uint64_t>(g.n_groups);
    p.pair_groups = static_cast<std::uint64_t>(g.n_pairs);
    p.pair_group_quartets = p.pair_groups * (p.pair_groups + 1) / 2;
    p.packed_pair_count = static_cast<std::uint64_t>(g.M) * (g.M + 1) / 2;
    p.packed_value_count = p.packed_pair_count * (p.packed_pair_count + 1) / 2;
    p.primitive_pairs = static_cast<std::uint64_t>(g.n_prim_pairs);
    return p;
}

}  // namespace

// ------------------------------------------------------------------ Boys --
namespace tlno::core::math {

// Orders 0..TLNO_GPU_BOYS_MMAX (the reference accepts any order; 128 is far
// beyond what an integral of g shells needs).
void require_order(const char* who, int n) {
    if (n < 0 || n > TLNO_GPU_BOYS_MMAX)
        throw std::runtime_error(std::string(who) + ": order must be in [0, " + std::to_string(TLNO_GPU_BOYS_MMAX) +
                                 "].");
}

double boys_single(int order, double x) {
    require_order("boys_single", order);
    std::vector<double> F(static_cast<std::size_t>(order) + 1);
    check(tlno_gpu_boys(1, &x, order, F.data()));
    return F[order];
}

void boys_r000_all(int max_n, double x, double p, double* out) {
    require_order("boys_r000_all", max_n);
    std::vector<double> F(static_cast<std::size_t>(max_n) + 1);
    check(tlno_gpu_boys(1, &x, max_n, F.data()));
    double s = 1.0;
    for (int n = 0; n <= max_n; ++n, s *= -2.0 * p) out[n] = s * F[n];
}

template <int W>
void boys_r000_all_batch(int max_n, const double* xs, double p, double* out) {
    boys_r000_all_batch_masked<W>(max_n, xs, p, out, (1u << W) - 1u);
}

// Inactive lanes may hold anything (NaN) and xs may be null when no lane is
// active (benchmarks/one_electron_engine/src/boys_mask_contract.cpp:17-47):
// only active lanes are read and sent to the device, inactive outputs are 0.
template <int W>
void boys_r000_all_batch_masked(int max_n, const double* xs, double p, double* out, unsigned active_mask) {
    static_assert(W == 1 || W == 2 || W == 4 || W == 8, "W must be 1, 2, 4 or 8");
    require_order("boys_r000_all", max_n);
    active_mask &= (1u << W) - 1u;
    double live[W];
    int nlive = 0;
    for (int lane = 0; lane < W; ++lane)
        if ((active_mask >> lane) & 1u) live[nlive++] = xs[lane];
    std::vector<double> F(static_cast<std::size_t>(W) * (max_n + 1));
    if (nlive > 0) check(tlno_gpu_boys(nlive, live, max_n, F.data()));
    for (int lane = 0, k = 0; lane < W; ++lane) {
        const bool on = (active_mask >> lane) & 1u;
        double s = 1.0;
        for (int n = 0; n <= max_n; ++n, s *= -2.0 * p) out[n * W + lane] = on ? s * F[k * (max_n + 1) + n] : 0.0;
        if (on) ++k;
    }
}

void boys_r000_all_x4(int max_n, const double* x4, double p, double* out) {
    boys_r000_all_batch<4>(max_n, x4, p, out);
}

template void boys_r000_all_batch<1>(int, const double*, double, double*);
template void boys_r000_all_batch<2>(int, const double*, double, double*);
template void boys_r000_all_batch<4>(int, const double*, double, double*);
template void boys_r000_all_batch<8>(int, const double*, double, double*);
template void boys_r000_all_batch_masked<1>(int, const double*, double, double*, unsigned);
template void boys_r000_all_batch_masked<2>(int, const double*, double, double*, unsigned);
template void boys_r000_all_batch_masked<4>(int, const double*, double, double*, unsigned);
template void boys_r000_all_batch_masked<8>(int, const double*, double, double*, unsigned);

}  // namespace tlno::core::math

// -------------------------------------------------------- twoe (packed) --
namespace tlno::integrals::twoe {

struct EriPackedPlanBackend {
    std::shared_ptr<GpuPlan> gpu;
};

struct EriPackedPlanBackendAccess {
    static const std::shared_ptr<const EriPackedPlanBackend>& get(const EriPackedPlan& p) { return p.backend_; }
    static void set(EriPackedPlan& p, std::shared_ptr<const EriPackedPlanBackend> b) { p.backend_ = std::move(b); }
};

namespace {
GpuPlan& gpu_of(const EriPackedPlan& plan) {
    const auto& b = EriPackedPlanBackendAccess::get(plan);
    if (!b || !b->gpu) throw std::runtime_error("eri_ao_tensor_hgp: null packed ERI plan.");
    return *b->gpu;
}

void check_schwarz(double v) {
    if (!finite_nonneg(v)) throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
}
}  // namespace

EriPackedPlan build_eri_packed_plan_for_G_rys(const std::vector<tlno::core::basis::Shell>& basis,
                                              const EriPackedBuildOptions& options) {
    check_schwarz(options.schwarz_threshold);
    auto b = std::make_shared<EriPackedPlanBackend>();
    b->gpu = make_gpu_plan(basis, 1);
    EriPackedPlan plan;
    plan.profile = profile_of(*b->gpu);
    EriPackedPlanBackendAccess::set(plan, std::move(b));
    return plan;
}

EriPackedGData build_eri_packed_for_G_rys(const EriPackedPlan& plan, const EriPackedBuildOptions& options) {
    check_schwarz(options.schwarz_threshold);
    require_schwarz_only(options.primitive_fastschwarz_threshold, options.decay_threshold);
    GpuPlan& g = gpu_of(plan);
    std::lock_guard<std::mutex> lk(g.mu);
    double v = 0.0;
    check(tlno_gpu_eri_checksum(g.h, options.schwarz_threshold, &v));
    return EriPackedGData{{v}};
}

EriPackedGData build_eri_packed_fingerprint_for_G_rys(const EriPackedPlan& plan,
                                                      const EriPackedBuildOptions& options) {
    check_schwarz(options.schwarz_threshold);
    require_schwarz_only(options.primitive_fastschwarz_threshold, options.decay_threshold);
    GpuPlan& g = gpu_of(plan);
    std::lock_guard<std::mutex> lk(g.mu);
    double f[7] = {};
    unsigned long long n = 0;
    check(tlno_gpu_eri_fingerprint(g.h, options.schwarz_threshold, f, &n));
    return EriPackedGData{{f[0], f[3], f[4], f[5], static_cast<double>(n)}};
}

EriDirectJKResult build_jk_direct_for_G_rys(const EriPackedPlan& plan, const std::vector<double>& density,
                                            const EriPackedBuildOptions& options) {
    check_schwarz(options.schwarz_threshold);
    check_heuristic_screens(options.primitive_fastschwarz_threshold, options.decay_threshold);
    if (options.density_threshold > 0.0 && options.decay_threshold > 0.0)
        throw std::runtime_error(
            "eri_ao_tensor_hgp: density screening requires Schwarz bounds and cannot be combined with the heuristic "
            "decay metric.");
    GpuPlan& g = gpu_of(plan);
    const std::size_t n = static_cast<std::size_t>(g.M) * g.M;
    std::vector<double> J(n), K(n);
    build_jk(g, density,
             {options.schwarz_threshold, options.density_threshold, options.primitive_fastschwarz_threshold,
              options.decay_threshold},
             J.data(), K.data());
    EriDirectJKResult r;
    r.M = g.M;
    r.g.resize(n);
    for (std::size_t i = 0; i < n; ++i) r.g[i] = J[i] - 0.5 * K[i];
    return r;
}

// ---------------------------------------------------- incremental state --
// The state is a three-step machine over the cached (D, G) pair: skip (D
// unchanged), exact (G(D) from scratch) or incremental (G += G(D - D_prev),
// screened by the delta-density threshold). Options are checked once by a
// rule table; each update scans D, picks its step, builds into locals and
// commits by swapping, so a throw leaves the state as it was
// (direct_build.ipp:562-763 for the caller-visible rules and messages).
namespace {

using IncOpts = EriDirectJKIncrementalOptions;

struct IncRule {
    bool (*holds)(const IncOpts&);
    const char* message;
};

bool delta_screen_on(const IncOpts& o) {
    return o.delta_density_threshold > 0.0 ||
           (o.use_first_delta_density_threshold && o.first_delta_density_threshold > 0.0);
}

// first failing rule names the error
const IncRule kIncRules[] = {
    {[](const IncOpts& o) { return o.exact_rebuild_interval != 0u; },
     "incremental exact_rebuild_interval must be >= 1."},
    {[](const IncOpts& o) { return finite_nonneg(o.base_options.schwarz_threshold); },
     "incremental schwarz_threshold must be finite and >= 0."},
    {[](const IncOpts& o) { return finite_nonneg(o.base_options.primitive_fastschwarz_threshold); },
     "incremental primitive_fastschwarz_threshold must be finite and >= 0."},
    {[](const IncOpts& o) { return finite_nonneg(o.base_options.decay_threshold); },
     "incremental decay_threshold must be finite and >= 0."},
    {[](const IncOpts& o) { return finite_nonneg(o.base_options.density_threshold); },
     "incremental base density_threshold must be finite and >= 0."},
    {[](const IncOpts& o) { return o.base_options.schwarz_threshold == 0.0 || o.base_options.decay_threshold == 0.0; },
     "incremental base options cannot combine Schwarz and decay screening."},
    {[](const IncOpts& o) { return finite_nonneg(o.delta_density_threshold); },
     "incremental delta_density_threshold must be finite and >= 0."},
    {[](const IncOpts& o) { return finite_nonneg(o.first_delta_density_threshold); },
     "incremental first_delta_density_threshold must be finite and >= 0."},
    {[](const IncOpts& o) { return !(delta_screen_on(o) && o.base_options.decay_threshold > 0.0); },
     "incremental delta-density screening cannot be combined with the heuristic decay metric."},
};

enum class IncStep { skip, exact, incremental };

// what one step does to the status and the counters
struct IncStepEffect {
    EriDirectJKUpdateStatus status;
    bool counts_exact, counts_incremental, counts_skipped;
};

IncStepEffect effect_of(IncStep s) {
    switch (s) {
        case IncStep::skip: return {EriDirectJKUpdateStatus::skipped_unchanged, false, false, true};
        case IncStep::exact: return {EriDirectJKUpdateStatus::exact_rebuild, true, false, false};
        default: return {EriDirectJKUpdateStatus::incremental_update, false, true, false};
    }
}

// an unchanged density skips unless forced; the first update, a forced one
// and every exact_rebuild_interval-th since the last exact one rebuild
IncStep choose_step(bool have_state, bool unchanged, bool forced, std::uint64_t since_exact,
                    std::uint64_t interval) {
    if (have_state && unchanged && !forced) return IncStep::skip;
    if (!have_state || forced || since_exact + 1u >= interval) return IncStep::exact;
    return IncStep::incremental;
}

// D - D_prev (or D itself before the first update), its largest magnitude and
// whether it is identically zero
struct DensityScan {
    std::vector<double> delta;
    double max_abs = 0.0;
    bool unchanged = false;
};

DensityScan scan_density(const std::vector<double>& d, const std::vector<double>& prev, bool have_prev) {
    DensityScan r;
    r.unchanged = have_prev;
    if (have_prev) r.delta.resize(d.size());
    for (std::size_t i = 0; i < d.size(); ++i) {
        if (!std::isfinite(d[i]))
            throw std::runtime_error("eri_ao_tensor_hgp: incremental density contains a non-finite value.");
        const double x = have_prev ? d[i] - prev[i] : d[i];
        if (!std::isfinite(x)) throw std::runtime_error("eri_ao_tensor_hgp: incremental density delta overflowed.");
        if (have_prev) {
            r.delta[i] = x;
            if (x != 0.0) r.unchanged = false;
        }
        r.max_abs = std::max(r.max_abs, std::abs(x));
    }
    return r;
}

void require_finite(const std::vector<double>& v, const char* message) {
    if (!std::all_of(v.begin(), v.end(), [](double x) { return std::isfinite(x); }))
        throw std::runtime_error(message);
}

}  // namespace

EriDirectJKIncrementalState::EriDirectJKIncrementalState(const EriPackedPlan& plan,
                                                         const EriDirectJKIncrementalOptions& options)
    : plan_(plan), options_(options) {
    M_ = gpu_of(plan_).M;
    for (const IncRule& r : kIncRules)
        if (!r.holds(options_)) throw std::runtime_error(std::string("eri_ao_tensor_hgp: ") + r.message);
    // an exact rebuild never screens with the full density
    options_.base_options.density_threshold = 0.0;
}

EriDirectJKUpdateStatus EriDirectJKIncrementalState::update(const std::vector<double>& density,
                                                            bool force_exact_rebuild) {
    const std::size_t n = static_cast<std::size_t>(M_) * M_;
    if (density.size() != n)
        throw std::runtime_error("eri_ao_tensor_hgp: incremental density must be spherical M x M row-major.");
    DensityScan scan = scan_density(density, density_, initialized_);
    const IncStep step = choose_step(initialized_, scan.unchanged, force_exact_rebuild,
                                     updates_since_exact_rebuild_, options_.exact_rebuild_interval);

    double threshold = 0.0;
    std::vector<double> next_g, next_density;
    if (step != IncStep::skip) {
        EriPackedBuildOptions bo = options_.base_options;
        if (step == IncStep::incremental) {
            const bool first = options_.use_first_delta_density_threshold && updates_since_exact_rebuild_ == 0u;
            threshold = first ? options_.first_delta_density_threshold : options_.delta_density_threshold;
            bo.density_threshold = threshold;
        }
        next_density = density;
        EriDirectJKResult built =
            build_jk_direct_for_G_rys(plan_, step == IncStep::exact ? density : scan.delta, bo);
        if (built.M != M_ || built.g.size() != n)
            throw std::runtime_error(
                "eri_ao_tensor_hgp: incremental direct J/K builder returned an inconsistent shape.");
        if (step == IncStep::exact) {
            require_finite(built.g, "eri_ao_tensor_hgp: exact direct J/K rebuild produced a non-finite value.");
            next_g = std::move(built.g);
        } else {
            require_finite(built.g, "eri_ao_tensor_hgp: incremental dG build produced a non-finite value.");
            next_g = g_;
            for (std::size_t i = 0; i < n; ++i) next_g[i] += built.g[i];
            require_finite(next_g, "eri_ao_tensor_hgp: incremental cached G overflowed.");
        }
    }

    // commit: nothing below throws
    const IncStepEffect fx = effect_of(step);
    if (step != IncStep::skip) {
        density_.swap(next_density);
        g_.swap(next_g);
        initialized_ = true;
        updates_since_exact_rebuild_ = fx.counts_exact ? 0u : updates_since_exact_rebuild_ + 1u;
    }
    ++counters_.updates;
    counters_.exact_rebuilds += fx.counts_exact;
    counters_.incremental_updates += fx.counts_incremental;
    counters_.skipped_updates += fx.counts_skipped;
    last_status_ = fx.status;
    last_max_abs_density_delta_ = step == IncStep::skip ? 0.0 : scan.max_abs;
    last_effective_density_threshold_ = threshold;
    return last_status_;
}

void EriDirectJKIncrementalState::reset() noexcept {
    density_.clear();
    g_.clear();
    counters_ = {};
    last_status_ = EriDirectJKUpdateStatus::uninitialized;
    last_max_abs_density_delta_ = 0.0;
    last_effective_density_threshold_ = 0.0;
    updates_since_exact_rebuild_ = 0;
    initialized_ = false;
}

}  // namespace tlno::integrals::twoe

// ------------------------------------------------------------ twoe::fock --
namespace tlno::integrals::twoe::fock {

struct DirectPlanBackend {
    std::shared_ptr<GpuPlan> gpu;
};

struct DirectPlanBackendAccess {
    static const std::shared_ptr<const DirectPlanBackend>& get(const DirectPlan& p) { return p.backend_; }
    static void set(DirectPlan& p, std::shared_ptr<const DirectPlanBackend> b) { p.backend_ = std::move(b); }
};

namespace {
GpuPlan& gpu_of(const DirectPlan& plan) {
    const auto& b = DirectPlanBackendAccess::get(plan);
    if (!b || !b->gpu) throw std::runtime_error("jk_direct: null J/K plan backend.");
    return *b->gpu;
}

void validate(const ScreeningOptions& o) {
    if (!finite_nonneg(o.integral_threshold))
        throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
    if (!finite_nonneg(o.density_threshold))
        throw std::runtime_error("eri_ao_tensor_hgp: density_threshold must be finite and >= 0.");
    if (o.density_threshold > 0.0 && o.decay_threshold > 0.0)
        throw std::runtime_error(
            "eri_ao_tensor_hgp: density screening requires Schwarz bounds and cannot be combined with the heuristic "
            "decay metric.");
    check_heuristic_screens(o.primitive_fastschwarz_threshold, o.decay_threshold);
}

Screens screens_of(const ScreeningOptions& o) {
    return {o.integral_threshold, o.density_threshold, o.primitive_fastschwarz_threshold, o.decay_threshold};
}
}  // namespace

ScreeningOptions screening_from_eri_options(const tlno::integrals::twoe::EriPackedBuildOptions& o) noexcept {
    return ScreeningOptions{o.schwarz_threshold, o.primitive_fastschwarz_threshold, o.decay_threshold,
                            o.density_threshold};
}

tlno::integrals::twoe::EriPackedBuildOptions screening_to_eri_options(const ScreeningOptions& o) noexcept {
    tlno::integrals::twoe::EriPackedBuildOptions e;
    e.schwarz_threshold = o.integral_threshold;
    e.primitive_fastschwarz_threshold = o.primitive_fastschwarz_threshold;
    e.decay_threshold = o.decay_threshold;
    e.density_threshold = o.density_threshold;
    return e;
}

DirectPlan build_direct_plan(const std::vector<tlno::core::basis::Shell>& basis, const ScreeningOptions& options) {
    validate(options);
    auto b = std::make_shared<DirectPlanBackend>();
    b->gpu = make_gpu_plan(basis, 1);
    DirectPlan plan;
    plan.profile = profile_of(*b->gpu);
    DirectPlanBackendAccess::set(plan, std::move(b));
    return plan;
}

DirectPlan adopt_eri_plan_for_transition(const tlno::integrals::twoe::EriPackedPlan& eri) {
    const auto& eb = tlno::integrals::twoe::EriPackedPlanBackendAccess::get(eri);
    if (!eb || !eb->gpu) throw std::runtime_error("eri_ao_tensor_hgp: null packed ERI plan.");
    auto b = std::make_shared<DirectPlanBackend>();
    b->gpu = eb->gpu;  // shares the immutable GPU plan
    DirectPlan plan;
    plan.profile = eri.profile;
    DirectPlanBackendAccess::set(plan, std::move(b));
    return plan;
}

CoulombJResult build_coulomb_j(const DirectPlan& plan, const std::vector<double>& density,
                               const ScreeningOptions& options) {
    validate(options);
    GpuPlan& g = gpu_of(plan);
    CoulombJResult r;
    r.M = g.M;
    r.j.resize(static_cast<std::size_t>(g.M) * g.M);
    build_jk(g, density, screens_of(options), r.j.data(), nullptr);
    return r;
}

ExchangeKResult build_exchange_k(const DirectPlan& plan, const std::vector<double>& density,
                                 const ScreeningOptions& options) {
    validate(options);
    GpuPlan& g = gpu_of(plan);
    ExchangeKResult r;
    r.M = g.M;
    r.k.resize(static_cast<std::size_t>(g.M) * g.M);
    build_jk(g, density, screens_of(options), nullptr, r.k.data());
    return r;
}

DirectResult build_rhf_g(const DirectPlan& plan, const std::vector<double>& density, const ScreeningOptions& options) {
    validate(options);
    GpuPlan& g = gpu_of(plan);
    const std::size_t n = static_cast<std::size_t>(g.M) * g.M;
    std::vector<double> J(n), K(n);
    build_jk(g, density, screens_of(options), J.data(), K.data());
    DirectResult r;
    r.M = g.M;
    r.g.resize(n);
    for (std::size_t i = 0; i < n; ++i) r.g[i] = J[i] - 0.5 * K[i];
    return r;
}

}  // namespace tlno::integrals::twoe::fock

