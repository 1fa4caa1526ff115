// Drives the facade's block-level one-electron API the way the reference's
// matrix driver does (integrals/one_electron/onee_matrices.cpp: groups, one
// ShellPairPrim per group pair, contracted blocks scattered into S/T/V), and
// the matrix API on the same basis, for comparison.
//
//   onee_block_driver input.txt output.txt
//
// input: n_comp, then per component "cx cy cz lx ly lz nprim e0 c0 ...";
// n_nuc, then "Z x y z" lines.  Screening is switched off (the block API
// evaluates every pair; the matrix API then does too).  Every block is
// accumulated twice into a zeroed buffer and halved: the block kernels add
// into the caller's block.  Output lines "name count v0 v1 ...".
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "core/basis/types.hpp"
#include "integrals/one_electron/onee_integrals.hpp"
#include "integrals/one_electron/onee_matrices.hpp"

namespace basis = tlno::core::basis;
namespace onee = tlno::integrals::onee;

static void put(FILE* f, const char* name, const std::vector<double>& v) {
    std::fprintf(f, "%s %zu", name, v.size());
    for (double x : v) std::fprintf(f, " %.17g", x);
    std::fprintf(f, "\n");
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    try {
        std::ifstream in(argv[1]);
        basis::BasisBuildResult build;
        int n = 0;
        in >> n;
        for (int i = 0; i < n; ++i) {
            basis::Shell s;
            int np = 0;
            in >> s.center[0] >> s.center[1] >> s.center[2] >> s.ang[0] >> s.ang[1] >> s.ang[2] >> np;
            for (int k = 0; k < np; ++k) {
                basis::Primitive p{};
                in >> p.exponent >> p.coefficient;
                s.primitives.push_back(p);
            }
            build.basis.push_back(s);
        }
        int nn = 0;
        in >> nn;
        for (int i = 0; i < nn; ++i) {
            int z = 0;
            std::array<double, 3> r{};
            in >> z >> r[0] >> r[1] >> r[2];
            build.nuclear_charges.push_back(z);
            build.nuclear_centers.push_back(r);
        }
        if (!in) throw std::runtime_error("unreadable input");
        onee::set_onee_screening_enabled(false);
        const onee::OneElectronPlan plan = onee::build_one_electron_plan(build, true);
        std::vector<double> S(static_cast<std::size_t>(n) * n, 0.0), T(S), V(S);
        for (std::size_t k = 0; k < plan.group_pairs.size(); ++k) {
            const auto& gp = plan.group_pairs[k];
            const onee::ShellGroup& A = plan.groups[gp.ga];
            const onee::ShellGroup& B = plan.groups[gp.gb];
            onee::ShellPairPrim spp = onee::build_shell_pair_prim(A, B);
            onee::cache_shell_pair_nucleus_hermite(A, B, spp);
            const std::size_t nb = A.basis_indices.size() * B.basis_indices.size();
            std::vector<double> s(nb, 0.0), t(nb, 0.0), v(nb, 0.0);
            for (int rep = 0; rep < 2; ++rep) {
                onee::contracted_shell_block_overlap_kinetic(A, B, spp, s.data(), t.data(), 0.0);
                onee::contracted_shell_block_nuclear(A, B, spp, v.data(), build.nuclear_charges,
                                                     build.nuclear_centers);
            }
            for (std::size_t i = 0; i < A.basis_indices.size(); ++i)
                for (std::size_t j = 0; j < B.basis_indices.size(); ++j) {
                    const std::size_t o = i * B.basis_indices.size() + j;
                    const int a = A.basis_indices[i], b = B.basis_indices[j];
                    for (auto* M : {&S, &T, &V}) {
                        const double x = 0.5 * (M == &S ? s[o] : M == &T ? t[o] : v[o]);
                        (*M)[static_cast<std::size_t>(a) * n + b] = x;
                        (*M)[static_cast<std::size_t>(b) * n + a] = x;
                    }
                }
        }
        const onee::OneElectronMatrices m = onee::evaluate_one_electron_matrices(build, plan);
        FILE* f = std::fopen(argv[2], "w");
        put(f, "S_blocks", S);
        put(f, "T_blocks", T);
        put(f, "V_blocks", V);
        put(f, "S", m.overlap.values);
        put(f, "T", m.kinetic.values);
        put(f, "V", m.nuclear.values);
        put(f, "plan", {double(plan.groups.size()), double(plan.group_pairs.size()), double(plan.pair_prims.size()),
                        double(plan.n_sph)});
        std::fclose(f);
        onee::set_onee_screening_enabled(true);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
