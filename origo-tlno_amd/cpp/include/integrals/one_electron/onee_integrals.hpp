// Block-level one-electron API of the C++ facade, with the names, types and
// signatures of the reference's integrals/one_electron/onee_integrals.hpp:
// 10-105.  Shell groups and primitive-pair data are host plan data, built as
// the reference builds them; the contracted S/T/V blocks are evaluated by the
// GPU engine and accumulated into the caller's blocks (the reference's block
// kernels accumulate and do not clear, benchmarks/one_electron_engine/src/
// transfer_contract.cpp:121-123).
#ifndef TLNO_AMD_INTEGRALS_ONE_ELECTRON_ONEE_INTEGRALS_HPP
#define TLNO_AMD_INTEGRALS_ONE_ELECTRON_ONEE_INTEGRALS_HPP

#include <array>
#include <vector>

#include "core/basis/types.hpp"

namespace tlno::integrals::onee {

// Cartesian components of one centre, angular momentum and exponent set,
// each with its own (normalized) coefficients.
struct ShellGroup {
    std::array<double, 3> center{};
    int L = 0;
    std::vector<int> basis_indices;
    std::vector<std::array<int, 3>> ang;
    std::vector<double> exponents;
    std::vector<std::vector<double>> coeffs;
};

// Primitive-pair data of one group pair, k = ia * n_prim_b + ib, every
// section n_prim_a * n_prim_b long inside one arena (storage).  Move-only:
// the named pointers point into storage.
struct ShellPairPrim {
    int n_prim_a = 0;
    int n_prim_b = 0;
    double ABx = 0.0, ABy = 0.0, ABz = 0.0, AB2 = 0.0;
    std::vector<double> storage;
    double* p = nullptr;              // alpha + beta
    double* inv2p = nullptr;          // 1 / (2p)
    double* chi = nullptr;            // alpha beta / p
    double* sss = nullptr;            // (pi/p)^{3/2} exp(-chi AB^2)
    double* expmChiAB2 = nullptr;     // exp(-chi AB^2)
    double* alpha = nullptr;
    double* beta = nullptr;
    double* PAx = nullptr;
    double* PAy = nullptr;
    double* PAz = nullptr;
    double* Px = nullptr;
    double* Py = nullptr;
    double* Pz = nullptr;
    double* e0x = nullptr;            // exp(-chi AB_axis^2)
    double* e0y = nullptr;
    double* e0z = nullptr;
    double* neg2pi_over_p = nullptr;  // -2 pi / p
    double* screen = nullptr;         // (2 pi/p) exp(-chi AB^2) max|c_a| max|c_b|
    // Nuclear Hermite projector cache of the reference's CPU kernels; the GPU
    // engine evaluates V by Rys quadrature and leaves it empty (an empty cache
    // means "not cached" in the reference as well).
    std::vector<double> nuclear_hermite;

    ShellPairPrim() = default;
    ShellPairPrim(ShellPairPrim&&) = default;
    ShellPairPrim& operator=(ShellPairPrim&&) = default;
    ShellPairPrim(const ShellPairPrim&) = delete;
    ShellPairPrim& operator=(const ShellPairPrim&) = delete;
};

ShellPairPrim build_shell_pair_prim(const ShellGroup& group_a, const ShellGroup& group_b);
void cache_shell_pair_nucleus_hermite(const ShellGroup& group_a, const ShellGroup& group_b, ShellPairPrim& spp);

// overlap_block / kinetic_block [n_a x n_b] (components in group order) +=
// the contracted S / T of the group pair; min_abs_overlap must be 0 (the
// reference's only value).
void contracted_shell_block_overlap_kinetic(const ShellGroup& group_a, const ShellGroup& group_b,
                                            const ShellPairPrim& spp, double* overlap_block,
                                            double* kinetic_block, double min_abs_overlap);

// nuclear_block [n_a x n_b] += V of the group pair over the given nuclei,
// primitive pairs screened at get_onee_nuclear_screen_eps() / sum|Z|.
void contracted_shell_block_nuclear(const ShellGroup& group_a, const ShellGroup& group_b, const ShellPairPrim& spp,
                                    double* nuclear_block, const std::vector<int>& nuclear_charges,
                                    const std::vector<std::array<double, 3>>& nuclear_centers);

constexpr double kNuclearScreenEps = 1.0e-15;

void set_onee_nuclear_screen_eps(double eps);
double get_onee_nuclear_screen_eps();

}  // namespace tlno::integrals::onee

#endif
