// Test infrastructure only (not shipped, not measured as the product).
//
// Fingerprint translation unit over the reference's Rys engine, built by
// oracle/ref.mk into oracle/_ref/libtlno_ref_fp_avx512.so (fixture
// generation only).  It turns on the reference's own validation switch
// TLNO_VALIDATE_RYS_STREAM_FINGERPRINT, under which the production packed
// endpoint emits every canonical value individually and returns
//   {checksum, indexed_moment_1, indexed_moment_2, indexed_moment_3, count}
// (rys_engine.ipp:87-105, 270-300; native_quartets.ipp:759-765).  The
// reference's production host (twoe_integrals.cpp) refuses the switch, so
// this file hosts the engine itself, as that host does: one include of
// rys_engine.ipp, which holds the plan builder, the packed endpoint and the
// direct J/K build.
//
// Exposed through extern "C":
//   ref_fp_stream(basis..., schwarz, out[5])
#define TLNO_VALIDATE_RYS_STREAM_FINGERPRINT 1
#include "integrals/two_electron/rys/rys_engine.ipp"

#include <exception>
#include <string>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string g_fp_err;

std::vector<tlno::core::basis::Shell> fp_basis(int n, const double* centers, const int* ang, const int* nprim,
                                               const double* exps, const double* coefs) {
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

}  // namespace

REF_API const char* ref_fp_last_error() { return g_fp_err.c_str(); }

// The reference's stream fingerprint of the canonical spherical ERIs at a
// Schwarz threshold: out = {checksum, m1, m2, m3, count}.
REF_API int ref_fp_stream(int n, const double* centers, const int* ang, const int* nprim, const double* exps,
                          const double* coefs, double schwarz, double* out) {
    try {
        namespace twoe = tlno::integrals::twoe;
        auto basis = fp_basis(n, centers, ang, nprim, exps, coefs);
        twoe::EriPackedBuildOptions opt;
        opt.schwarz_threshold = schwarz;
        auto plan = twoe::build_eri_packed_plan_for_G_rys(basis, opt);
        auto data = twoe::build_eri_packed_for_G_rys(plan, opt);
        if (data.values.size() != 5) {
            g_fp_err = "fingerprint build returned " + std::to_string(data.values.size()) + " values";
            return 2;
        }
        for (int k = 0; k < 5; ++k) out[k] = data.values[static_cast<std::size_t>(k)];
        return 0;
    } catch (const std::exception& ex) {
        g_fp_err = ex.what();
        return 1;
    }
}
