// One-electron S, T, V of the C++ facade, with the names and signatures of
// the reference's integrals/one_electron/onee_matrices.hpp:10-79.  The plan
// owns a GPU plan; Cartesian output is in component order, spherical output
// in the plan's spherical AO order.
#ifndef TLNO_AMD_INTEGRALS_ONE_ELECTRON_ONEE_MATRICES_HPP
#define TLNO_AMD_INTEGRALS_ONE_ELECTRON_ONEE_MATRICES_HPP

#include <memory>

#include "core/basis/types.hpp"

namespace tlno::integrals::onee {

namespace basis = tlno::core::basis;

struct OneElectronMatrices {
    basis::DenseMatrix overlap;
    basis::DenseMatrix kinetic;
    basis::DenseMatrix nuclear;
    explicit OneElectronMatrices(int n = 0) : overlap(n), kinetic(n), nuclear(n) {}
};

struct OneElectronPlanBackend;

struct OneElectronPlan {
    int n_cart = 0;
    int n_sph = 0;
    std::shared_ptr<const OneElectronPlanBackend> backend;
};

OneElectronPlan build_one_electron_plan(const basis::BasisBuildResult& build);
OneElectronPlan build_one_electron_plan(const basis::BasisBuildResult& build, bool cache_nuclear_hermite);
OneElectronPlan build_one_electron_spherical_plan(const basis::BasisBuildResult& build,
                                                  bool cache_nuclear_hermite = true);
OneElectronMatrices build_one_electron_spherical_matrices(const basis::BasisBuildResult& build);
OneElectronMatrices evaluate_one_electron_matrices(const basis::BasisBuildResult& build,
                                                   const OneElectronPlan& plan);
void evaluate_one_electron_matrices_into(const basis::BasisBuildResult& build, const OneElectronPlan& plan,
                                         OneElectronMatrices& mats);
void evaluate_one_electron_matrices_sph_into(const basis::BasisBuildResult& build, const OneElectronPlan& plan,
                                             OneElectronMatrices& mats);
OneElectronMatrices build_one_electron_matrices(const basis::BasisBuildResult& build);

// The GPU engine evaluates every group pair (no decay screen): these keep the
// reference's switches and report them, without changing results.
void set_onee_screening_enabled(bool enabled);
bool onee_screening_enabled();
void set_onee_group_pair_decay_eps(double eps);
double onee_group_pair_decay_eps();

}  // namespace tlno::integrals::onee

#endif
