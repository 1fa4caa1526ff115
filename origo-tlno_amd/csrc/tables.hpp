// Host-side generation of the Boys and Rys interpolation tables used by the
// device code.  Everything is computed at library initialisation from the
// defining integrals; nothing is tabulated in the source.
//
//   Boys:  F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt, sampled on T_k = k/32,
//          k = 0..BOYS_NGRID-1, m = 0..BOYS_MMAX, by the positive series
//          e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) in long double.
//          Device evaluation: 7-term Taylor expansion about the nearest grid
//          point using dF_m/dT = -F_{m+1} (|T - T_k| <= 1/64, remainder
//          <= 64^-7 / 7! < 5e-17 relative); orders up to BOYS_MMAX - 6.
//
//   Rys:   roots t_i^2 and weights w_i of the rule exact for
//          int_0^1 f(t^2) exp(-T t^2) dt, deg f <= 2N-1 (the reference's
//          definition, integrals/two_electron/rys/rys_engine.ipp:3714-3719).
//          Generated by the discretized Stieltjes procedure on a 320-point
//          Gauss-Legendre grid followed by the Golub-Welsch eigen-solve, in
//          long double.  Stored as degree-11 polynomials in u = T - k - 1/2 on
//          unit intervals [k, k+1) up to RYS_TTAB(N); above it the half-range
//          Hermite asymptote r_i^2/T, w_i/sqrt(T) (rys_engine.ipp:3543-3547)
//          is exact to binary64.  Maximum relative error of the polynomial
//          pieces is about 2e-15 (checked in tests against the oracle).
#pragma once

#include <vector>

namespace tlno_amd {

constexpr int BOYS_MMAX = 24;          // orders stored per grid point
constexpr int BOYS_STEPS = 32;         // grid points per unit of T
constexpr int BOYS_NGRID = 41 * BOYS_STEPS + 1; // T in [0, 41], step 1/32
constexpr double BOYS_TCUT = 40.0;     // asymptotic + upward recursion beyond
constexpr int BOYS_TAYLOR = 7;

// The coarse copy the lane kernels with one or two roots stage in LDS:
// T_k = k/8, orders 0..11, 9-term Taylor expansion (|T - T_k| <= 1/16,
// remainder <= 16^-9 / 9! < 2e-16 relative) for F_0..F_3; 31.6 KB.
constexpr int BOYS_LDS_STEPS = 8;
constexpr int BOYS_LDS_NGRID = 41 * BOYS_LDS_STEPS + 1;
constexpr int BOYS_LDS_M = 12;
constexpr int BOYS_LDS_TAYLOR = 9;

constexpr int RYS_NMAX = 9;   // shells up to g: at most 9 roots
constexpr int RYS_DEG = 11;             // polynomial degree per interval
constexpr int RYS_NCOEF = RYS_DEG + 1;

// first T handled by the asymptote, per N (index N)
inline int rys_ttab(int N) {
    static const int t[10] = {0, 41, 49, 55, 61, 67, 72, 77, 83, 88};
    return t[N];
}

struct HostTables {
    std::vector<double> boys;        // [BOYS_NGRID][BOYS_MMAX+1]
    std::vector<double> boys_lds;    // [BOYS_LDS_NGRID][BOYS_LDS_M]
    std::vector<double> rys;         // concatenated per N: [nint][2N][RYS_NCOEF]
    int rys_off[RYS_NMAX + 2] = {0}; // offset of N's block in rys
    double herm_r2[RYS_NMAX + 1][RYS_NMAX] = {};
    double herm_w[RYS_NMAX + 1][RYS_NMAX] = {};
};

// Builds (once, thread-safe) and returns the tables for N <= nmax_req.
const HostTables& host_tables();

// Accurate host evaluation used by the generator (also exported for tests).
void rys_accurate(int N, double T, double* roots, double* weights);
double boys_accurate(int m, double T);

}  // namespace tlno_amd
