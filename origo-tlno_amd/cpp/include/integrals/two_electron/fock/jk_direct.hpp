// Direct J / K / G endpoints of the C++ facade, with the names and signatures
// of the reference's integrals/two_electron/fock/jk_direct.hpp:18-96.
#ifndef TLNO_AMD_INTEGRALS_TWO_ELECTRON_FOCK_JK_DIRECT_HPP
#define TLNO_AMD_INTEGRALS_TWO_ELECTRON_FOCK_JK_DIRECT_HPP

#include <memory>
#include <vector>

#include "integrals/two_electron/fock/eri_direct_jk.hpp"

namespace tlno::integrals::twoe::fock {

struct ScreeningOptions {
    double integral_threshold = tlno::integrals::twoe::kEriPackedDefaultSchwarzThreshold;
    double primitive_fastschwarz_threshold = 0.0;
    double decay_threshold = 0.0;
    double density_threshold = 0.0;
};

ScreeningOptions screening_from_eri_options(const tlno::integrals::twoe::EriPackedBuildOptions& options) noexcept;
tlno::integrals::twoe::EriPackedBuildOptions screening_to_eri_options(const ScreeningOptions& options) noexcept;

struct DirectPlanBackend;

struct DirectPlan {
    tlno::integrals::twoe::EriPackedPlanProfile profile;

private:
    std::shared_ptr<const DirectPlanBackend> backend_;
    friend struct DirectPlanBackendAccess;
};

DirectPlan build_direct_plan(const std::vector<tlno::core::basis::Shell>& basis, const ScreeningOptions& options = {});
DirectPlan adopt_eri_plan_for_transition(const tlno::integrals::twoe::EriPackedPlan& plan);

struct DirectResult {
    std::vector<double> g;
    int M = 0;
};

struct CoulombJResult {
    std::vector<double> j;
    int M = 0;
};

struct ExchangeKResult {
    std::vector<double> k;
    int M = 0;
};

CoulombJResult build_coulomb_j(const DirectPlan& plan, const std::vector<double>& density,
                               const ScreeningOptions& options);
// K[a,c] = sum_(b,d) P[b,d] (ab|cd), for a real symmetric density.
ExchangeKResult build_exchange_k(const DirectPlan& plan, const std::vector<double>& density,
                                 const ScreeningOptions& options);
DirectResult build_rhf_g(const DirectPlan& plan, const std::vector<double>& density,
                         const ScreeningOptions& options);

}  // namespace tlno::integrals::twoe::fock

#endif
