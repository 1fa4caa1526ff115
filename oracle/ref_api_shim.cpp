// Test infrastructure only (not shipped, not measured as the product).
//
// extern "C" driver over the reference's PUBLIC C++ API, compiled together
// with the reference sources under /root/reference by oracle/ref.mk into
// oracle/_ref/libtlno_ref.so.  It lets the Python golden-vector generator and
// the CPU-baseline leg of bench.py call the reference engine through ctypes.
//
// Interfaces driven (paths relative to the reference root):
//   core/shell/shell_contract.hpp:10-17   normalize_contracted_primitives
//   core/math/general_math.hpp:6-22       boys_single, boys_r000_all
//   integrals/two_electron/fock/jk_direct.hpp:54-96
//       build_direct_plan, build_coulomb_j, build_exchange_k, build_rhf_g
//   integrals/two_electron/twoe_integrals.hpp:104-113
//       build_eri_packed_plan_for_G_rys, build_eri_packed_for_G_rys (checksum)
//   integrals/one_electron/onee_matrices.hpp:31-79
//       build_one_electron_matrices, build_one_electron_spherical_matrices
//
// Basis arguments are Cartesian components exactly like the reference's
// std::vector<Shell> (core/basis/types.hpp:31-36): component i has centre
// centers[3i..3i+2], angular powers ang[3i..3i+2], nprim[i] primitives whose
// exponents/coefficients are consecutive in exps/coefs.  Coefficients are the
// reference's stored (already normalized) coefficients.
#include "core/basis/types.hpp"
#include "core/math/general_math.hpp"
#include "core/shell/shell_contract.hpp"
#include "integrals/one_electron/onee_matrices.hpp"
#include "integrals/two_electron/fock/jk_direct.hpp"
#include "integrals/two_electron/twoe_integrals.hpp"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

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
        for (int p = 0; p < nprim[i]; ++p, ++off) {
            s.primitives.push_back({exps[off], coefs[off]});
        }
    }
    return basis;
}

}  // namespace

REF_API const char* ref_last_error() { return g_err.c_str(); }

REF_API void ref_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

REF_API int ref_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// normalize_contracted_primitives: raw (exps, coefs) of one component -> stored coefs.
REF_API int ref_normalize(int nprim, const double* exps, const double* coefs,
                          const int* ang, double* out_coefs) {
    try {
        std::vector<double> e(exps, exps + nprim), c(coefs, coefs + nprim);
        auto prims = tlno::core::shell::normalize_contracted_primitives(
            e, c, {ang[0], ang[1], ang[2]});
        for (int i = 0; i < nprim; ++i) out_coefs[i] = prims[static_cast<std::size_t>(i)].coefficient;
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

REF_API double ref_boys_single(int order, double x) {
    return tlno::core::math::boys_single(order, x);
}

REF_API void ref_boys_r000_all(int max_n, double x, double p, double* out) {
    tlno::core::math::boys_r000_all(max_n, x, p, out);
}

// Spherical function count M of the reference's direct plan.
REF_API int ref_plan_M(int n, const double* centers, const int* ang,
                       const int* nprim, const double* exps, const double* coefs) {
    try {
        auto basis = make_basis(n, centers, ang, nprim, exps, coefs);
        tlno::integrals::twoe::fock::ScreeningOptions opt;
        auto plan = tlno::integrals::twoe::fock::build_direct_plan(basis, opt);
        return plan.profile.M;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// which: 0 = J (build_coulomb_j), 1 = K (build_exchange_k), 2 = G (build_rhf_g).
// density/out are M x M row-major in the spherical AO basis.
// repeats > 1 runs the build that many times on one plan (timing use);
// elapsed_ns (optional) receives the wall time of those builds only.
REF_API int ref_direct_build_ex(int n, const double* centers, const int* ang,
                                const int* nprim, const double* exps,
                                const double* coefs, double schwarz,
                                double density_threshold, double fastschwarz,
                                double decay, int which,
                                const double* density, int M, double* out,
                                int repeats, long long* elapsed_ns);

REF_API int ref_direct_build(int n, const double* centers, const int* ang,
                             const int* nprim, const double* exps,
                             const double* coefs, double schwarz,
                             double density_threshold, int which,
                             const double* density, int M, double* out,
                             int repeats, long long* elapsed_ns) {
    return ref_direct_build_ex(n, centers, ang, nprim, exps, coefs, schwarz, density_threshold, 0.0, 0.0, which,
                               density, M, out, repeats, elapsed_ns);
}

// The same build with the primitive FASTSCHWARZ and decay thresholds of
// ScreeningOptions (integrals/two_electron/fock/jk_direct.hpp:18-24).
REF_API int ref_direct_build_ex(int n, const double* centers, const int* ang,
                                const int* nprim, const double* exps,
                                const double* coefs, double schwarz,
                                double density_threshold, double fastschwarz,
                                double decay, int which,
                                const double* density, int M, double* out,
                                int repeats, long long* elapsed_ns) {
    try {
        namespace fock = tlno::integrals::twoe::fock;
        auto basis = make_basis(n, centers, ang, nprim, exps, coefs);
        fock::ScreeningOptions opt;
        opt.integral_threshold = schwarz;
        opt.density_threshold = density_threshold;
        opt.primitive_fastschwarz_threshold = fastschwarz;
        opt.decay_threshold = decay;
        auto plan = fock::build_direct_plan(basis, opt);
        if (plan.profile.M != M) {
            g_err = "ref_direct_build: M mismatch";
            return 2;
        }
        std::vector<double> d(density, density + static_cast<std::size_t>(M) * M);
        std::vector<double> result;
        if (repeats < 1) repeats = 1;
        timespec t0{}, t1{};
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (int r = 0; r < repeats; ++r) {
            if (which == 0) result = fock::build_coulomb_j(plan, d, opt).j;
            else if (which == 1) result = fock::build_exchange_k(plan, d, opt).k;
            else result = fock::build_rhf_g(plan, d, opt).g;
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        if (elapsed_ns)
            *elapsed_ns = (t1.tv_sec - t0.tv_sec) * 1000000000LL + (t1.tv_nsec - t0.tv_nsec);
        std::memcpy(out, result.data(), result.size() * sizeof(double));
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// Production checksum stream (sum of all canonical spherical ERIs).
REF_API int ref_eri_checksum(int n, const double* centers, const int* ang,
                             const int* nprim, const double* exps,
                             const double* coefs, double schwarz, double* checksum) {
    try {
        namespace twoe = tlno::integrals::twoe;
        auto basis = make_basis(n, centers, ang, nprim, exps, coefs);
        twoe::EriPackedBuildOptions opt;
        opt.schwarz_threshold = schwarz;
        auto plan = twoe::build_eri_packed_plan_for_G_rys(basis, opt);
        auto data = twoe::build_eri_packed_for_G_rys(plan, opt);
        *checksum = data.values.empty() ? 0.0 : data.values[0];
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// S, T, V.  spherical = 0: Cartesian n x n (component order);
// spherical = 1: spherical M x M.  nuc_R is 3*n_nuc bohr coordinates.
REF_API int ref_onee(int n, const double* centers, const int* ang,
                     const int* nprim, const double* exps, const double* coefs,
                     int n_nuc, const int* nuc_Z, const double* nuc_R,
                     int spherical, int screening, double* S, double* T,
                     double* V, int* n_out) {
    try {
        namespace onee = tlno::integrals::onee;
        tlno::core::basis::BasisBuildResult build;
        build.basis = make_basis(n, centers, ang, nprim, exps, coefs);
        for (int k = 0; k < n_nuc; ++k) {
            build.nuclear_charges.push_back(nuc_Z[k]);
            build.nuclear_centers.push_back({nuc_R[3 * k], nuc_R[3 * k + 1], nuc_R[3 * k + 2]});
        }
        onee::set_onee_screening_enabled(screening != 0);
        onee::OneElectronMatrices m =
            spherical ? onee::build_one_electron_spherical_matrices(build)
                      : onee::build_one_electron_matrices(build);
        onee::set_onee_screening_enabled(true);
        const int nn = m.overlap.size;
        if (n_out) *n_out = nn;
        const std::size_t bytes = static_cast<std::size_t>(nn) * nn * sizeof(double);
        if (S) std::memcpy(S, m.overlap.values.data(), bytes);
        if (T) std::memcpy(T, m.kinetic.values.data(), bytes);
        if (V) std::memcpy(V, m.nuclear.values.data(), bytes);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}
