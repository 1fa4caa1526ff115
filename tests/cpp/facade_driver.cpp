// Exercises the C++ facade (the reference's engine API over the GPU engine)
// the way a reference caller would, reading a basis and two densities from a
// text file and writing every result as "name count v0 v1 ..." lines.
//
//   facade_driver input.txt output.txt
//
// input: n_comp, then per component "cx cy cz lx ly lz nprim e0 c0 e1 c1 ...";
// n_nuc, then "Z x y z" lines; M, then two M x M spherical densities.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/basis/types.hpp"
#include "core/math/general_math.hpp"
#include "integrals/one_electron/onee_matrices.hpp"
#include "integrals/two_electron/fock/eri_direct_jk.hpp"
#include "integrals/two_electron/fock/jk_direct.hpp"
#include "integrals/two_electron/twoe_integrals.hpp"

namespace basis = tlno::core::basis;
namespace twoe = tlno::integrals::twoe;
namespace fock = tlno::integrals::twoe::fock;
namespace onee = tlno::integrals::onee;

static void put(FILE* f, const char* name, const std::vector<double>& v) {
    std::fprintf(f, "%s %zu", name, v.size());
    for (double x : v) std::fprintf(f, " %.17g", x);
    std::fprintf(f, "\n");
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input output\n", argv[0]);
        return 2;
    }
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
        int M = 0;
        in >> M;
        std::vector<double> D0(static_cast<std::size_t>(M) * M), D1(D0.size());
        for (double& x : D0) in >> x;
        for (double& x : D1) in >> x;
        if (!in) throw std::runtime_error("facade_driver: malformed input");

        FILE* out = std::fopen(argv[2], "w");
        if (!out) throw std::runtime_error("facade_driver: cannot open output");
        // fock:: endpoints
        fock::ScreeningOptions so;
        const fock::DirectPlan plan = fock::build_direct_plan(build.basis, so);
        put(out, "M", {double(plan.profile.M), double(plan.profile.shell_groups), double(plan.profile.pair_groups)});
        put(out, "J", fock::build_coulomb_j(plan, D0, so).j);
        put(out, "K", fock::build_exchange_k(plan, D0, so).k);
        put(out, "G", fock::build_rhf_g(plan, D0, so).g);
        // twoe:: packed plan, stream checksum, RHF G
        twoe::EriPackedBuildOptions eo;
        const twoe::EriPackedPlan eplan = twoe::build_eri_packed_plan_for_G_rys(build.basis, eo);
        put(out, "checksum", twoe::build_eri_packed_for_G_rys(eplan, eo).values);
        put(out, "fingerprint", twoe::build_eri_packed_fingerprint_for_G_rys(eplan, eo).values);
        put(out, "G_twoe", twoe::build_jk_direct_for_G_rys(eplan, D0, eo).g);
        const fock::DirectPlan adopted = fock::adopt_eri_plan_for_transition(eplan);
        put(out, "G_adopted", fock::build_rhf_g(adopted, D1, so).g);
        // incremental SCF state: exact, incremental, skipped, forced exact
        twoe::EriDirectJKIncrementalOptions io;
        io.base_options = eo;
        io.exact_rebuild_interval = 3;
        twoe::EriDirectJKIncrementalState state(eplan, io);
        std::vector<double> status;
        status.push_back(double(state.update(D0)));
        status.push_back(double(state.update(D1)));
        put(out, "G_incremental", state.g());
        status.push_back(double(state.update(D1)));
        status.push_back(double(state.update(D0)));
        status.push_back(double(state.update(D1)));  // third change: periodic exact rebuild
        status.push_back(double(state.update(D1, true)));
        put(out, "status", status);
        put(out, "counters", {double(state.counters().updates), double(state.counters().exact_rebuilds),
                              double(state.counters().incremental_updates), double(state.counters().skipped_updates)});
        put(out, "G_final", state.g());
        // errors keep the reference's messages
        try {
            twoe::EriPackedBuildOptions bad;
            bad.schwarz_threshold = -1.0;
            (void)twoe::build_jk_direct_for_G_rys(eplan, D0, bad);
            put(out, "error_raised", {0.0});
        } catch (const std::runtime_error& e) {
            std::fprintf(out, "error %s\n", e.what());
        }
        // incremental options: the first broken rule names the error; a
        // failed update leaves the state untouched
        {
            twoe::EriDirectJKIncrementalOptions zero = io, mixed = io;
            zero.exact_rebuild_interval = 0;
            mixed.delta_density_threshold = 1e-9;
            mixed.base_options.schwarz_threshold = 0.0;
            mixed.base_options.decay_threshold = 1e-6;
            int raised = 0;
            for (const auto* o : {&zero, &mixed}) {
                try {
                    twoe::EriDirectJKIncrementalState bad_state(eplan, *o);
                } catch (const std::runtime_error& e) {
                    ++raised;
                    std::fprintf(out, "inc_error %s\n", e.what());
                }
            }
            const std::uint64_t before = state.counters().updates;
            try {
                (void)state.update(std::vector<double>(3, 0.0));
            } catch (const std::runtime_error&) {
                ++raised;
            }
            const bool kept = state.counters().updates == before &&
                              state.last_status() == twoe::EriDirectJKUpdateStatus::exact_rebuild;
            put(out, "inc_errors", {double(raised), kept ? 1.0 : 0.0});
        }
        // one-electron matrices
        const onee::OneElectronMatrices mc = onee::build_one_electron_matrices(build);
        put(out, "S_cart", mc.overlap.values);
        put(out, "T_cart", mc.kinetic.values);
        put(out, "V_cart", mc.nuclear.values);
        const onee::OneElectronMatrices ms = onee::build_one_electron_spherical_matrices(build);
        put(out, "S_sph", ms.overlap.values);
        put(out, "V_sph", ms.nuclear.values);
        // Boys seeds
        std::vector<double> r(5);
        tlno::core::math::boys_r000_all(4, 2.5, 0.7, r.data());
        r.push_back(tlno::core::math::boys_single(3, 17.25));
        put(out, "boys", r);
        std::fclose(out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
