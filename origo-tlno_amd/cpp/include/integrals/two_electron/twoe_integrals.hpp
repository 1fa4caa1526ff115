// Two-electron plan and canonical-stream API of the C++ facade, with the
// names and signatures of the reference's
// integrals/two_electron/twoe_integrals.hpp:13-113.  The plan owns a GPU plan
// (C ABI tlno_gpu_plan) in the pure-spherical AO basis.
#ifndef TLNO_AMD_INTEGRALS_TWO_ELECTRON_TWOE_INTEGRALS_HPP
#define TLNO_AMD_INTEGRALS_TWO_ELECTRON_TWOE_INTEGRALS_HPP

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "core/basis/types.hpp"

namespace tlno::integrals::twoe {

// Canonical (a >= b) pair and (ab >= cd) quartet ids of the packed stream.
struct EriPackedGAccessor {
    static constexpr std::uint32_t pair_index(int a, int b) noexcept {
        const std::uint64_t hi = static_cast<std::uint64_t>(a < b ? b : a);
        const std::uint64_t lo = static_cast<std::uint64_t>(a < b ? a : b);
        return static_cast<std::uint32_t>(hi * (hi + 1u) / 2u + lo);
    }
    static constexpr std::uint64_t quartet_index(std::uint32_t ab, std::uint32_t cd) noexcept {
        const std::uint64_t hi = ab >= cd ? ab : cd;
        const std::uint64_t lo = ab >= cd ? cd : ab;
        return hi * (hi + 1u) / 2u + lo;
    }
};

struct EriPackedGData {
    std::vector<double> values;  // {checksum of the canonical stream}
};

inline constexpr double kEriPackedDefaultSchwarzThreshold = 1.0e-12;

struct EriPackedBuildOptions {
    double schwarz_threshold = kEriPackedDefaultSchwarzThreshold;
    double primitive_fastschwarz_threshold = 0.0;  // J/K builds; 0 for the stream endpoints
    double decay_threshold = 0.0;                  // J/K builds; 0 for the stream endpoints
    double density_threshold = 0.0;                // J/K builds only
};

struct EriPackedPlanClassStats {
    int la = 0, lb = 0, lc = 0, ld = 0;
    std::uint64_t pair_group_quartets = 0;
    std::uint64_t primitive_quartets = 0;
    std::uint64_t lhs_component_refs = 0;
    std::uint64_t rhs_component_refs = 0;
};

struct EriPackedPlanProfile {
    int M = 0;
    std::uint64_t shell_groups = 0;
    std::uint64_t pair_groups = 0;
    std::uint64_t pair_group_quartets = 0;
    std::uint64_t packed_pair_count = 0;
    std::uint64_t packed_value_count = 0;
    std::uint64_t primitive_pairs = 0;
    std::uint64_t max_primitive_pairs_per_pair_group = 0;
    std::vector<EriPackedPlanClassStats> class_stats;
};

struct EriPackedPlanBackend;

struct EriPackedPlan {
    EriPackedPlanProfile profile;

private:
    std::shared_ptr<const EriPackedPlanBackend> backend_;
    friend struct EriPackedPlanBackendAccess;
};

EriPackedPlan build_eri_packed_plan_for_G_rys(const std::vector<tlno::core::basis::Shell>& basis,
                                              const EriPackedBuildOptions& options);

EriPackedGData build_eri_packed_for_G_rys(const EriPackedPlan& plan, const EriPackedBuildOptions& options);

// What build_eri_packed_for_G_rys returns in the reference's
// TLNO_VALIDATE_RYS_STREAM_FINGERPRINT build (native_quartets.ipp:759-765):
// values = {checksum, indexed_moment_1, indexed_moment_2, indexed_moment_3,
// canonical_count}, the indexed moments weighting each canonical value by
// hashes of its quartet index (rys_engine.ipp:87-105).  Here a separate
// entry point rather than a build switch.
EriPackedGData build_eri_packed_fingerprint_for_G_rys(const EriPackedPlan& plan,
                                                      const EriPackedBuildOptions& options);

}  // namespace tlno::integrals::twoe

#endif
