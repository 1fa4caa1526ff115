// Test infrastructure only (not shipped, not measured as the product).
//
// Audit translation unit over the reference's Rys engine, built by
// oracle/ref.mk into oracle/_ref/libtlno_ref_audit.so.  It follows the
// reference's own experiment-host pattern
// (benchmarks/two_electron_engine/src/rys_native_lane_ablation.cpp:8-9):
// define TLNO_BUILD_RYS_AUDIT and include the engine, so the internal
// routines are reachable from this file.  Exposed through extern "C":
//   - shell grouping            build_groups_h           rys_engine.ipp:1966-2062
//   - raw Cartesian ERI blocks  compute_quartet_block_rys_impl rys_engine.ipp:13754-13842
//   - Rys roots / weights       rys_roots_weights (table), rys_roots_weights_accurate
//                               rys_engine.ipp:3720-3837
//   - 2e Boys                   boys_eval_hgp            rys_engine.ipp:554-580
//   - production spherical ERI stream with a recording Emit policy
//                               build_eri_packed_for_G_rys_spherical_stream_impl
//                               rys_engine.ipp:18229-18280 (native W8 route on AVX-512)
#define TLNO_BUILD_RYS_AUDIT 1
#include "integrals/two_electron/rys/rys_engine.ipp"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace tlno::integrals::twoe {
namespace audit_shim {

thread_local std::string g_err;

std::vector<tlno::core::basis::Shell> make_basis(int n, const double* centers,
                                                 const int* ang, const int* nprim,
                                                 const double* exps,
                                                 const double* coefs) {
    std::vector<tlno::core::basis::Shell> basis(static_cast<std::size_t>(n));
    std::size_t off = 0;
    for (int i = 0; i < n; ++i) {
        auto& s = basis[static_cast<std::size_t>(i)];
        s.label = "c" + std::to_string(i);
        for (int k = 0; k < 3; ++k) {
            s.center[k] = centers[3 * i + k];
            s.ang[k] = ang[3 * i + k];
        }
        for (int p = 0; p < nprim[i]; ++p, ++off) s.primitives.push_back({exps[off], coefs[off]});
    }
    return basis;
}

struct Recorded {
    std::vector<std::uint32_t> lhs, rhs;
    std::vector<double> val;
};
thread_local Recorded* g_rec = nullptr;

struct RecordingEmit {
    static inline void emit(std::vector<double>&, std::uint32_t lhs_pair,
                            std::uint32_t rhs_pair, double value) noexcept {
        g_rec->lhs.push_back(lhs_pair);
        g_rec->rhs.push_back(rhs_pair);
        g_rec->val.push_back(value);
    }
};

thread_local Recorded g_last_stream;

}  // namespace audit_shim
}  // namespace tlno::integrals::twoe

using namespace tlno::integrals::twoe;
using namespace tlno::integrals::twoe::audit_shim;

REF_API const char* ref_audit_last_error() { return g_err.c_str(); }

// Groups as the reference forms them: per group L, number of components,
// and the basis index of each component (max 15 per group).
REF_API int ref_audit_groups(int n, const double* centers, const int* ang,
                             const int* nprim, const double* exps,
                             const double* coefs, int max_groups, int* n_groups,
                             int* group_L, int* group_comp /* [max_groups*15] */) {
    try {
        auto basis = make_basis(n, centers, ang, nprim, exps, coefs);
        auto groups = build_groups_h(basis);
        *n_groups = static_cast<int>(groups.size());
        if (*n_groups > max_groups) { g_err = "too many groups"; return 2; }
        for (int g = 0; g < *n_groups; ++g) {
            group_L[g] = groups[g].L;
            for (int c = 0; c < 15; ++c)
                group_comp[g * 15 + c] =
                    c < static_cast<int>(groups[g].basis_indices.size()) ? groups[g].basis_indices[c] : -1;
        }
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// Unit-diagonal-normalized Cartesian block (ga gb | gc gd), laid out
// [comp_a][comp_b][comp_c][comp_d] in each group's component order.
// compute_quartet_block_rys_impl returns the block in descending-lex local
// Cartesian order before the N_A factors; the shim maps to component order
// and applies n_ang of each component.
REF_API int ref_audit_cart_block(int n, const double* centers, const int* ang,
                                 const int* nprim, const double* exps,
                                 const double* coefs, int ga, int gb, int gc,
                                 int gd, double* out) {
    try {
        auto basis = make_basis(n, centers, ang, nprim, exps, coefs);
        auto groups = build_groups_h(basis);
        const int ng = static_cast<int>(groups.size());
        if (ga < 0 || gb < 0 || gc < 0 || gd < 0 || ga >= ng || gb >= ng || gc >= ng || gd >= ng) {
            g_err = "group index out of range";
            return 2;
        }
        const auto& A = groups[ga];
        const auto& B = groups[gb];
        const auto& C = groups[gc];
        const auto& D = groups[gd];
        std::vector<double> blk;
        compute_quartet_block_rys_impl(A, B, C, D, nullptr, nullptr, nullptr, nullptr, blk);
        const int na = n_cart(A.L), nb = n_cart(B.L), nc = n_cart(C.L), nd = n_cart(D.L);
        auto local = [](const ContractedShellGroupH& g, int comp) {
            const auto& t = g.angular[static_cast<std::size_t>(comp)];
            return triple_idx_within(g.L, t[0], t[1]);
        };
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b)
                for (int c = 0; c < nc; ++c)
                    for (int d = 0; d < nd; ++d) {
                        const std::size_t src =
                            ((static_cast<std::size_t>(local(A, a)) * nb + local(B, b)) * nc + local(C, c)) * nd +
                            local(D, d);
                        const double norm = A.n_ang[a] * B.n_ang[b] * C.n_ang[c] * D.n_ang[d];
                        out[((static_cast<std::size_t>(a) * nb + b) * nc + c) * nd + d] = blk[src] * norm;
                    }
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// mode 0: production table route rys_roots_weights; 1: rys_roots_weights_accurate.
REF_API void ref_audit_rys(int N, double T, int mode, double* roots, double* weights) {
    if (mode == 0) rys_roots_weights(N, T, roots, weights);
    else rys_roots_weights_accurate(N, T, roots, weights);
}

REF_API double ref_audit_boys_hgp(int m, double x) { return boys_eval_hgp(m, x); }

// Run the production spherical stream with a recording Emit policy; returns the
// number of canonical values. Fetch them with ref_audit_stream_fetch.
REF_API long long ref_audit_sph_stream(int n, const double* centers, const int* ang,
                                       const int* nprim, const double* exps,
                                       const double* coefs, double schwarz) {
    try {
        auto basis = make_basis(n, centers, ang, nprim, exps, coefs);
        auto bd = build_basis_data_h(basis);
        EriPackedBuildOptions opt;
        opt.schwarz_threshold = schwarz;
        g_last_stream = Recorded{};
        g_rec = &g_last_stream;
        build_eri_packed_for_G_rys_spherical_stream_impl<RecordingEmit>(*bd, opt);
        g_rec = nullptr;
        return static_cast<long long>(g_last_stream.val.size());
    } catch (const std::exception& ex) {
        g_err = ex.what();
        g_rec = nullptr;
        return -1;
    }
}

REF_API void ref_audit_stream_fetch(unsigned* lhs, unsigned* rhs, double* val) {
    const std::size_t n = g_last_stream.val.size();
    std::memcpy(lhs, g_last_stream.lhs.data(), n * sizeof(unsigned));
    std::memcpy(rhs, g_last_stream.rhs.data(), n * sizeof(unsigned));
    std::memcpy(val, g_last_stream.val.data(), n * sizeof(double));
}
