// One-electron S, T, V of the C++ facade, with the names, types and
// signatures of the reference's integrals/one_electron/onee_matrices.hpp:
// 10-79.  The plan holds the reference's host plan data (shell groups,
// screened group pairs, primitive-pair data, spherical layout) plus the GPU
// plan that evaluates the matrices; Cartesian output is in component order,
// spherical output in the reference's spherical AO order.
#ifndef TLNO_AMD_INTEGRALS_ONE_ELECTRON_ONEE_MATRICES_HPP
#define TLNO_AMD_INTEGRALS_ONE_ELECTRON_ONEE_MATRICES_HPP

#include <cstddef>
#include <memory>
#include <vector>

#include "core/basis/types.hpp"
#include "integrals/one_electron/onee_integrals.hpp"

namespace tlno::integrals::onee {

namespace basis = tlno::core::basis;

struct OneElectronMatrices {
    basis::DenseMatrix overlap;
    basis::DenseMatrix kinetic;
    basis::DenseMatrix nuclear;
    explicit OneElectronMatrices(int n = 0) : overlap(n), kinetic(n), nuclear(n) {}
};

struct OneElectronGroupPair {
    int ga = 0;
    int gb = 0;
    bool do_st = true;
    bool do_v = true;
};

// GPU side of a plan (no reference counterpart).
struct OneElectronPlanBackend;

struct OneElectronPlan {
    std::vector<ShellGroup> groups;
    std::vector<OneElectronGroupPair> group_pairs;
    std::vector<ShellPairPrim> pair_prims;  // parallel to group_pairs
    std::size_t nuclear_hermite_cache_bytes = 0;
    int n_sph = 0;
    std::vector<std::vector<int>> sph_base;  // spherical AO base of each Cartesian shell of a group
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

// Screening switches of the reference (group-pair decay estimate, taken at
// plan build; nuclear primitive-pair estimate, taken at evaluation): the GPU
// engine applies the same rules.
void set_onee_screening_enabled(bool enabled);
bool onee_screening_enabled();
void set_onee_group_pair_decay_eps(double eps);
double onee_group_pair_decay_eps();

}  // namespace tlno::integrals::onee

#endif
