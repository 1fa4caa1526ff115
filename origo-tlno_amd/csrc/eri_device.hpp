// Device side of the Rys-quadrature ERI engine for CDNA4 (gfx950).
//
// One wave (64 lanes) owns one task: a fixed bra shell pair P and up to 64
// consecutive ket shell pairs Q of one class, Schwarz-sorted so that the
// surviving kets of P form a prefix.  Each lane evaluates one contracted
// shell quartet (P|Q) entirely in registers (a task with at most 32 kets
// gives each ket several lanes, which split its bra primitive pairs):
//
//   for bra primitive pair x ket primitive pair:
//     T = rho |PQ|^2, Rys roots/weights (Boys for N = 1, piecewise
//     polynomials + Hermite asymptote for N >= 2),
//     per root and axis: 2D VRR G[e][f], ket HRR, bra HRR (all unrolled at
//     compile time for the class), then acc[o] += Ix * Iy * (w Iz).
//
// and then digests its block against the density: the J contribution of the
// shared bra pair is reduced across the wave with cross-lane shuffles before
// one atomic per component; the ket-side J and the four exchange blocks go
// out as FP64 atomics into pair-space accumulators Jacc/Kacc, which a
// finalize kernel symmetrizes (J = Jacc + Jacc^T, K = Kacc + Kacc^T).
//
// Classes whose full Cartesian block would not fit in registers are split
// into compile-time "chunks": a fixed bra component AC and/or ket component
// CC per kernel instance, so a lane accumulates at most 108 outputs.
//
// The per-primitive-quartet formulas follow the reference's generic Rys path
// (integrals/two_electron/rys/rys_engine.ipp:13540-13628, HRR 3155-3363):
//   B00 = t^2/(2(p+q)), B10 = (1 - q t^2/(p+q))/(2p), B01 = (1 - p t^2/(p+q))/(2q),
//   C00 = (P-A) - q/(p+q) PQ t^2, D00 = (Q-C) + p/(p+q) PQ t^2,
//   prefactor 2 pi^{5/2} w_ab w_cd / sqrt(p+q), w = K c_a c_b / p.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include "tables.hpp"

namespace tlno_amd {

__host__ __device__ constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

template <int L>
struct CartT {
    int e[3][ncart(L)];
    double nang[ncart(L)];  // 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)
};

__host__ __device__ constexpr int dfact2m1(int n) {
    int v = 1;
    for (int k = 2 * n - 1; k > 0; k -= 2) v *= k;
    return v;
}

// constexpr 1/sqrt for the few values that occur ((2l-1)!! products, l <= 4)
__host__ __device__ constexpr double inv_sqrt_small(int v) {
    return v == 1 ? 1.0
         : v == 3 ? 0.57735026918962576451
         : v == 9 ? 1.0 / 3.0
         : v == 15 ? 0.25819888974716112568
         : v == 105 ? 0.09759000729485332666
         : v == 45 ? 0.14907119849998597976
         : v == 225 ? 1.0 / 15.0
         : 0.0;
}

template <int L>
__host__ __device__ constexpr CartT<L> make_cart() {
    CartT<L> t{};
    int k = 0;
    for (int x = L; x >= 0; --x)
        for (int y = L - x; y >= 0; --y) {
            t.e[0][k] = x;
            t.e[1][k] = y;
            t.e[2][k] = L - x - y;
            t.nang[k] = inv_sqrt_small(dfact2m1(x) * dfact2m1(y) * dfact2m1(L - x - y));
            ++k;
        }
    return t;
}

// ---------------------------------------------------------------- tables ---
struct DevTables {
    const double* boys;   // [BOYS_NGRID][BOYS_MMAX+1]
    const double* boys_lds;  // [BOYS_LDS_NGRID][BOYS_LDS_M], the copy kernels stage in LDS
    const double* rys;    // polynomial pieces, see tables.hpp
    int rys_off[RYS_NMAX + 2];
    double herm_r2[RYS_NMAX + 1][RYS_NMAX];
    double herm_w[RYS_NMAX + 1][RYS_NMAX];
};

// Reciprocal and reciprocal square root of a positive normal double: the
// hardware estimates (v_rcp_f64, v_rsq_f64) refined by two Newton-Raphson
// steps each, r <- r + r (1 - x r) and y <- y + (y/2)(1 - x y^2); the error
// of each step is about the square of the previous one, which leaves a few
// ulp.  They replace the IEEE-exact division and square root sequences in
// the per-primitive work (every argument there is a sum of Gaussian
// exponents or a Boys argument T >= 40).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return y;
}

// F_0..F_MMAX at T.  sb: the coarse table staged in LDS by the calling
// kernel (nullptr: the fine table in global memory)
template <int MMAX>
__device__ __forceinline__ void boys_all(const DevTables& tb, double T, double* F, const double* sb = nullptr) {
    static_assert(MMAX + BOYS_TAYLOR - 1 <= BOYS_MMAX, "Boys order beyond the tabulated orders");
    constexpr bool kLdsOrders = MMAX + BOYS_LDS_TAYLOR - 1 < BOYS_LDS_M;  // the LDS copy holds orders 0..11
    if (kLdsOrders && sb && T < BOYS_TCUT) {
        const int k = max(static_cast<int>(T * BOYS_LDS_STEPS + 0.5), 0);
        const double dd = k * (1.0 / BOYS_LDS_STEPS) - T;
        const double* row = sb + k * BOYS_LDS_M;
#pragma unroll
        for (int m = 0; m <= (kLdsOrders ? MMAX : 0); ++m) {
            double v = row[m + BOYS_LDS_TAYLOR - 1];
#pragma unroll
            for (int j = BOYS_LDS_TAYLOR - 1; j >= 1; --j) v = row[m + j - 1] + v * dd * (1.0 / j);
            F[m] = v;
        }
    } else if (T < BOYS_TCUT) {
        // T >= 0 for every integral argument; the clamp keeps a negative
        // argument from a caller inside the table
        const int k = max(static_cast<int>(T * BOYS_STEPS + 0.5), 0);
        const double dd = k * (1.0 / BOYS_STEPS) - T;  // F_m(T) = sum_j F_{m+j}(T_k) dd^j / j!
        const double* row = tb.boys + k * (BOYS_MMAX + 1);
#pragma unroll
        for (int m = 0; m <= MMAX; ++m) {
            double v = row[m + BOYS_TAYLOR - 1];
#pragma unroll
            for (int j = BOYS_TAYLOR - 1; j >= 1; --j) v = row[m + j - 1] + v * dd * (1.0 / j);
            F[m] = v;
        }
    } else {
        const double isT = rsqrt_nr(T);
        const double iT = isT * isT;
        double f = 0.88622692545275801365 * isT;  // sqrt(pi)/2
        F[0] = f;
        if (MMAX > 0) {
            const double ex = exp(-T);
            const double h = 0.5 * iT;
#pragma unroll
            for (int m = 0; m < MMAX; ++m) {
                f = ((2 * m + 1) * f - ex) * h;
                F[m + 1] = f;
            }
        }
    }
}

template <int N>
__device__ __forceinline__ void rys_roots(const DevTables& tb, double T, double* r, double* w,
                                          const double* sb = nullptr) {
    if constexpr (N == 1) {
        double F[2];
        boys_all<1>(tb, T, F, sb);
        r[0] = F[1] * rcp_nr(F[0]);
        w[0] = F[0];
    } else if constexpr (N == 2) {
        if (T < 49.0) {
            // two-point Gauss rule in x = t^2 from the moments F_0..F_3 of its
            // weight: x^2 + c1 x + c0 is orthogonal to 1 and x, its roots are
            // the nodes, and the weights match F_0 and F_1 (12 Boys table
            // values per T instead of 48 interpolation coefficients; the
            // gathers, not the arithmetic, bound the classes that use it)
            double F[4];
            boys_all<3>(tb, T, F, sb);
            const double id = rcp_nr(F[0] * F[2] - F[1] * F[1]);
            const double c1 = (F[1] * F[2] - F[0] * F[3]) * id;
            const double c0 = (F[1] * F[3] - F[2] * F[2]) * id;
            const double disc = c1 * c1 - 4.0 * c0;
            const double xb = 0.5 * (disc * rsqrt_nr(disc) - c1);  // c1 < 0: no cancellation
            const double xs = c0 * rcp_nr(xb);
            const double iw = rcp_nr(xb - xs);
            r[0] = xs;
            r[1] = xb;
            w[0] = (xb * F[0] - F[1]) * iw;
            w[1] = (F[1] - xs * F[0]) * iw;
        } else {
            const double isT = rsqrt_nr(T);
            const double iT = isT * isT;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                r[i] = tb.herm_r2[N][i] * iT;
                w[i] = tb.herm_w[N][i] * isT;
            }
        }
    } else {
        constexpr int ttab[10] = {0, 41, 49, 55, 61, 67, 72, 77, 83, 88};
        if (T < ttab[N]) {
            const int k = max(static_cast<int>(T), 0);
            const double u = T - k - 0.5;
            const double* c = tb.rys + tb.rys_off[N] + k * (2 * N * RYS_NCOEF);
#pragma unroll
            for (int v = 0; v < 2 * N; ++v) {
                const double* cv = c + v * RYS_NCOEF;
                double s = cv[RYS_NCOEF - 1];
#pragma unroll
                for (int p = RYS_NCOEF - 2; p >= 0; --p) s = fma(s, u, cv[p]);
                if (v < N) r[v] = s;
                else w[v - N] = s;
            }
        } else {
            const double isT = rsqrt_nr(T);
            const double iT = isT * isT;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                r[i] = tb.herm_r2[N][i] * iT;
                w[i] = tb.herm_w[N][i] * isT;
            }
        }
    }
}

// Rys roots and weights one root at a time (the team evaluator gives each
// lane a single root of a primitive quartet).
template <int N>
struct RootEval {
    const double* c = nullptr;  // table branch: this T's interval (nullptr: asymptotic branch)
    double u = 0.0, iT = 0.0, isT = 0.0;
    double F0 = 0.0, F1 = 0.0;  // N == 1
    __device__ __forceinline__ RootEval(const DevTables& tb, double T) {
        if constexpr (N == 1) {
            double F[2];
            boys_all<1>(tb, T, F);
            F0 = F[0];
            F1 = F[1];
        } else {
            constexpr int ttab[10] = {0, 41, 49, 55, 61, 67, 72, 77, 83, 88};
            if (T < ttab[N]) {
                const int k = max(static_cast<int>(T), 0);
                u = T - k - 0.5;
                c = tb.rys + tb.rys_off[N] + k * (2 * N * RYS_NCOEF);
            } else {
                isT = rsqrt_nr(T);
                iT = isT * isT;
            }
        }
    }
    __device__ __forceinline__ void get(const DevTables& tb, int r, double& t2, double& w) const {
        if constexpr (N == 1) {
            t2 = F1 * rcp_nr(F0);
            w = F0;
        } else {
            if (c) {
                const double* cr = c + r * RYS_NCOEF;
                const double* cw = c + (N + r) * RYS_NCOEF;
                double s = cr[RYS_NCOEF - 1], sw = cw[RYS_NCOEF - 1];
#pragma unroll
                for (int p = RYS_NCOEF - 2; p >= 0; --p) {
                    s = fma(s, u, cr[p]);
                    sw = fma(sw, u, cw[p]);
                }
                t2 = s;
                w = sw;
            } else {
                t2 = tb.herm_r2[N][r] * iT;
                w = tb.herm_w[N][r] * isT;
            }
        }
    }
};

// ------------------------------------------------------------ pair data ---
struct PairData {
    const double* A;        // [npair][3] centre of the bra-side (higher-L) shell
    const double* AB;       // [npair][3] A - B
    const int* ao;          // [npair][2] internal Cartesian AO base of A and B
    const int* poff;        // [npair] first primitive pair
    const int* pn;          // [npair] primitive pair count
    const double4* pP;      // [nprim] {Px, Py, Pz, p}
    const double* pw;       // [nprim] K c_a c_b / p
    const int* diag;        // [npair] 1 if A and B are the same shell
};

// ----------------------------------------------------- one shell quartet ---
// R is the arithmetic type of the 2D recurrences and the contraction: double,
// or float for the mixed-precision route (roots, weights, T and the
// prefactor are always formed in double and rounded once).
template <typename R, int LA, int LB, int LC, int LD>
__device__ __forceinline__ void axis_table(R C00, R D00, R B00, R B10, R B01, R g00, R AB, R CD,
                                           R (&I)[LA + 1][LB + 1][LC + 1][LD + 1]) {
    constexpr int LAB = LA + LB, LCD = LC + LD;
    R G[LAB + 1][LCD + 1];
    G[0][0] = g00;
    if constexpr (LAB > 0) G[1][0] = C00 * g00;
#pragma unroll
    for (int e = 1; e < LAB; ++e) G[e + 1][0] = C00 * G[e][0] + (R(e) * B10) * G[e - 1][0];
#pragma unroll
    for (int f = 0; f < LCD; ++f) {
#pragma unroll
        for (int e = 0; e <= LAB; ++e) {
            R v = D00 * G[e][f];
            if (e > 0) v = fma(R(e) * B00, G[e - 1][f], v);
            if (f > 0) v = fma(R(f) * B01, G[e][f - 1], v);
            G[e][f + 1] = v;
        }
    }
    // ket HRR: H[e][c][d], (c, d+1) = (c+1, d) + CD (c, d)
    R H[LAB + 1][LCD + 1][LD + 1];
#pragma unroll
    for (int e = 0; e <= LAB; ++e) {
#pragma unroll
        for (int c = 0; c <= LCD; ++c) H[e][c][0] = G[e][c];
#pragma unroll
        for (int d = 0; d < LD; ++d)
#pragma unroll
            for (int c = 0; c + d + 1 <= LCD; ++c) H[e][c][d + 1] = fma(CD, H[e][c][d], H[e][c + 1][d]);
    }
    // bra HRR: (a, b+1) = (a+1, b) + AB (a, b)
#pragma unroll
    for (int c = 0; c <= LC; ++c)
#pragma unroll
        for (int d = 0; d <= LD; ++d) {
            R Jb[LAB + 1][LB + 1];
#pragma unroll
            for (int e = 0; e <= LAB; ++e) Jb[e][0] = H[e][c][d];
#pragma unroll
            for (int b = 0; b < LB; ++b)
#pragma unroll
                for (int a = 0; a + b + 1 <= LAB; ++a) Jb[a][b + 1] = fma(AB, Jb[a][b], Jb[a + 1][b]);
#pragma unroll
            for (int a = 0; a <= LA; ++a)
#pragma unroll
                for (int b = 0; b <= LB; ++b) I[a][b][c][d] = Jb[a][b];
        }
}

// Output chunk of a class: all bra components a (AC < 0) or a in
// [AC, AC + AN); all ket components c (CC < 0) or c in [CC, CC + CN).
template <int LA, int LB, int LC, int LD, int AC, int AN, int CC, int CN>
struct Shape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int NR = (LAB + LCD) / 2 + 1;
    static constexpr int NAL = AC < 0 ? NA : AN;
    static constexpr int NCL = CC < 0 ? NC : CN;
    static constexpr int NOUT = NAL * NB * NCL * ND;
    __device__ static constexpr int acomp(int ai) { return AC < 0 ? ai : AC + ai; }
    __device__ static constexpr int ccomp(int ci) { return CC < 0 ? ci : CC + ci; }
};

constexpr double kTwoPi52 = 34.986836655249725693;  // 2 pi^{5/2}

// Classes whose per-primitive work is large enough to pay for per-lane
// primitive indexing walk the flattened (ket primitive, bra primitive) grid.
#ifndef TLNO_FLAT_MIN_L
#define TLNO_FLAT_MIN_L 4
#endif
__host__ __device__ constexpr bool flat_prims(int LA, int LB, int LC, int LD) {
    return LA + LB + LC + LD >= TLNO_FLAT_MIN_L;
}

// Accumulates the lane's block slice (unit-diagonal components) into acc over
// its share of the primitive quartets: bra primitive pairs [ib0, ib1) of P
// (a task's slice) times all ket primitive pairs of Q, of which the lane takes
// every js-th starting at j0 -- of the flattened grid (bra index fastest) for
// flat_prims classes, else of the bra pairs (with every ket pair).
// uniform: P, ib0 and ib1 are the same in every lane of the wave (lane 0
// holds the range when some lanes pass an empty one).
// FS: the primitive FASTSCHWARZ screen is on (fs_min > 0); a separate
// instance, so builds without the screen run the unscreened code unchanged.
template <typename R, int LA, int LB, int LC, int LD, int AC, int AN, int CC, int CN, bool FS = false>
__device__ __forceinline__ void quartet(const DevTables& tb, const PairData& pd, int P, int Q, int ib0, int ib1,
                                        int j0, int js, bool uniform,
                                        R (&acc)[Shape<LA, LB, LC, LD, AC, AN, CC, CN>::NOUT],
                                        double fs_min = 0.0, const double* sb = nullptr) {
    using S = Shape<LA, LB, LC, LD, AC, AN, CC, CN>;
    constexpr CartT<LA> ca = make_cart<LA>();
    constexpr CartT<LB> cb = make_cart<LB>();
    constexpr CartT<LC> cc = make_cart<LC>();
    constexpr CartT<LD> cd = make_cart<LD>();
#pragma unroll
    for (int o = 0; o < S::NOUT; ++o) acc[o] = R(0);

    const double Ax = pd.A[3 * P], Ay = pd.A[3 * P + 1], Az = pd.A[3 * P + 2];
    const R ABx = R(pd.AB[3 * P]), ABy = R(pd.AB[3 * P + 1]), ABz = R(pd.AB[3 * P + 2]);
    const double Cx = pd.A[3 * Q], Cy = pd.A[3 * Q + 1], Cz = pd.A[3 * Q + 2];
    const R CDx = R(pd.AB[3 * Q]), CDy = R(pd.AB[3 * Q + 1]), CDz = R(pd.AB[3 * Q + 2]);
    // the block is linear in the primitives, so the shares of a quartet held
    // by different lanes or waves digest independently
    const int pb0 = pd.poff[P];
    const int pk0 = pd.poff[Q], pkn = pd.pn[Q];

    auto prim = [&](const double4 Pp, const double wab, const double4 Qp, const double wcd) {
        const double q = Qp.w;
        const double p = Pp.w;
        // primitive FASTSCHWARZ screen (this lane's quartet only): a failing
        // quartet gets a zero prefactor (a mask, not a branch: the lanes of a
        // wave run the same primitive code);
        // (w_ab w_cd)^2 / sqrt(pq) >= fs_min  <=>  (w_ab w_cd)^4 >= fs_min^2 p q
        double keep = 1.0;
        if constexpr (FS) {
            const double w2 = wab * wcd, w4 = (w2 * w2) * (w2 * w2);
            keep = w4 < (fs_min * fs_min) * (p * q) ? 0.0 : 1.0;
        }
        const double pq = p + q;
        const double sipq = rsqrt_nr(pq);
        const double ipq = sipq * sipq;
        const double PQx = Pp.x - Qp.x, PQy = Pp.y - Qp.y, PQz = Pp.z - Qp.z;
        const double T = p * q * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
        const double pref = FS ? kTwoPi52 * wab * wcd * sipq * keep : kTwoPi52 * wab * wcd * sipq;
        double rt[S::NR], wt[S::NR];
        rys_roots<S::NR>(tb, T, rt, wt, sb);
        const R qo = R(q * ipq), po = R(p * ipq);
        const R hpq = R(0.5 * ipq), i2p = R(0.5 * rcp_nr(p)), i2q = R(0.5 * rcp_nr(q));
        const R PAx = R(Pp.x - Ax), PAy = R(Pp.y - Ay), PAz = R(Pp.z - Az);
        const R QCx = R(Qp.x - Cx), QCy = R(Qp.y - Cy), QCz = R(Qp.z - Cz);
        const R pqx = R(PQx), pqy = R(PQy), pqz = R(PQz);
#pragma unroll
        for (int r = 0; r < S::NR; ++r) {
            const R t2 = R(rt[r]);
            const R B00 = hpq * t2;
            const R B10 = i2p * (R(1) - qo * t2);
            const R B01 = i2q * (R(1) - po * t2);
            const R qt = qo * t2, pt = po * t2;
            R Ix[LA + 1][LB + 1][LC + 1][LD + 1];
            R Iy[LA + 1][LB + 1][LC + 1][LD + 1];
            R Iz[LA + 1][LB + 1][LC + 1][LD + 1];
            axis_table<R, LA, LB, LC, LD>(PAx - qt * pqx, QCx + pt * pqx, B00, B10, B01, R(1), ABx, CDx, Ix);
            axis_table<R, LA, LB, LC, LD>(PAy - qt * pqy, QCy + pt * pqy, B00, B10, B01, R(1), ABy, CDy, Iy);
            axis_table<R, LA, LB, LC, LD>(PAz - qt * pqz, QCz + pt * pqz, B00, B10, B01, R(pref * wt[r]), ABz,
                                          CDz, Iz);
            int o = 0;
#pragma unroll
            for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
                for (int b = 0; b < S::NB; ++b)
#pragma unroll
                    for (int ci = 0; ci < S::NCL; ++ci)
#pragma unroll
                        for (int d = 0; d < S::ND; ++d, ++o) {
                            const int a = S::acomp(ai), c = S::ccomp(ci);
                            const R xy = Ix[ca.e[0][a]][cb.e[0][b]][cc.e[0][c]][cd.e[0][d]] *
                                         Iy[ca.e[1][a]][cb.e[1][b]][cc.e[1][c]][cd.e[1][d]];
                            acc[o] = fma(xy, Iz[ca.e[2][a]][cb.e[2][b]][cc.e[2][c]][cd.e[2][d]], acc[o]);
                        }
        }
    };
    // the uniform walk serves the smaller flattened-grid classes too (same
    // visiting order with one lane per ket); for the larger ones the second
    // copy of the primitive code costs more than the scalar loads gain
    // (organic30: (dp|ps) 1.09 -> 1.01 ms, but (ds|pp) 0.85 -> 1.12 ms)
    constexpr bool kUniformWalk =
        !flat_prims(LA, LB, LC, LD) || (LA + LB + LC + LD <= 4 && S::NOUT * (LC + LD + 1) <= 108);
    if (kUniformWalk && uniform && js == 1) {
        // one lane per ket: every lane walks the same bra primitive pairs
        // (lane 0's range; lanes without a ket of their own run it on a
        // stand-in ket and are not digested), so their data are wave-uniform
        // and read through the scalar cache (the pair data are read-only for
        // the kernel's lifetime: the constant address space is exact)
        using CDouble = const __attribute__((address_space(4))) double;
        CDouble* cpP = (CDouble*)(pd.pP);
        CDouble* cpw = (CDouble*)(pd.pw);
        const int ub = __builtin_amdgcn_readfirstlane(pb0);
        const int u0 = __builtin_amdgcn_readfirstlane(ib0), u1 = __builtin_amdgcn_readfirstlane(ib1);
        for (int ik = 0; ik < pkn; ++ik) {
            const double4 Qp = pd.pP[pk0 + ik];
            const double wcd = pd.pw[pk0 + ik];
            for (int ib = u0; ib < u1; ++ib) {
                const long i4 = 4L * (ub + ib);
                prim(make_double4(cpP[i4], cpP[i4 + 1], cpP[i4 + 2], cpP[i4 + 3]), cpw[ub + ib], Qp, wcd);
            }
        }
    } else if constexpr (flat_prims(LA, LB, LC, LD)) {
        const int nb = ib1 - ib0;
        if (nb > 0) {
            const int ntot = pkn * nb;
            int ik = j0 / nb, ib = j0 - ik * nb;
            const int dk = js / nb, db = js - dk * nb;
            for (int j = j0; j < ntot; j += js) {
                prim(pd.pP[pb0 + ib0 + ib], pd.pw[pb0 + ib0 + ib], pd.pP[pk0 + ik], pd.pw[pk0 + ik]);
                ik += dk;
                ib += db;
                if (ib >= nb) {
                    ib -= nb;
                    ++ik;
                }
            }
        }
    } else {
        // ket primitives outside: each lane reads its ket primitive data once
        for (int ik = 0; ik < pkn; ++ik) {
            const double4 Qp = pd.pP[pk0 + ik];
            const double wcd = pd.pw[pk0 + ik];
            for (int ib = ib0 + j0; ib < ib1; ib += js) prim(pd.pP[pb0 + ib], pd.pw[pb0 + ib], Qp, wcd);
        }
    }
    // unit-diagonal normalization of each component
    int o = 0;
#pragma unroll
    for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
        for (int b = 0; b < S::NB; ++b)
#pragma unroll
            for (int ci = 0; ci < S::NCL; ++ci)
#pragma unroll
                for (int d = 0; d < S::ND; ++d, ++o) {
                    const int a = S::acomp(ai), c = S::ccomp(ci);
                    acc[o] *= R(ca.nang[a] * cb.nang[b] * cc.nang[c] * cd.nang[d]);
                }
}

// ------------------------------------------------------------ kernels ----
// Counters are spread over kCounterSlots slots of four (work item w adds to
// slot w % kCounterSlots), so the waves of a launch do not all serialize on
// the same four addresses; readers sum the slots.
constexpr int kCounterSlots = 64;

struct JKArgs {
    PairData pd;
    DevTables tb;
    const int4* tasks;  // range mode {P, q_begin, q_end, slice};
                        // list mode {P, offset bits 0-31, count | offset bits 32-55 << 8,
                        //            bra primitive slice ib0 | ib1 << 16 (split == 0)}
    int ntasks;
    int split;          // list mode: waves per task (bra primitive slices); 0: one wave per
                        // task, on the slice the task names
    const double* D;    // internal Cartesian density, M x M
    double* J;          // Jacc
    double* K;          // Kacc
    int M;
    int do_j, do_k;
    const int* cut;     // density screen: ket range of P ends at cut[6 P] (pointer offset by ket class; nullptr: off)
    const int* qlist;   // list mode: kets of a task are qlist[offset .. offset + count)
    unsigned long long* counter;  // [kCounterSlots][4]: [0] shell, [1] primitive quartets evaluated,
                                  // [2], [3] the FP32 share
    // checksum mode (cw != nullptr, no J/K): chk += sum over the block of
    // cw[pair kind of P][a][b] * cw[pair kind of Q][c][d] * (ab|cd), for
    // P != Q; pair kind = 2 * class + (1 if a diagonal pair)
    const double* cw;   // [12][36]
    double* chk;
    // molecule batches (pmol != nullptr): pair P belongs to molecule pmol[P],
    // whose D / J / K blocks start at mbase[mol] with row stride mM[mol]
    const int* pmol;
    const long long* mbase;
    const int* mM;
    // persistent launch (stride > 0): each wave takes work items
    // w, w + stride, ... below nwork (range mode: tasks; list mode: tasks x split)
    int stride;
    int nwork;
    // team evaluator: waves per shell quartet, each taking a slice of its
    // primitive combinations (0 or 1: one wave)
    int team_slices;
    // team evaluator, quartet mode (tquart != nullptr): work item i is the
    // shell quartet {P, Q} = tquart[i], i < ntquart (no task decoding, no
    // empty waves)
    const int2* tquart;
    long long ntquart;
    // primitive FASTSCHWARZ screen (0: off): a primitive quartet is left out
    // when (w_ab w_cd)^2 / sqrt(p q) < fs_min, i.e. fs_ab fs_cd < t^2 with
    // fs = 24.739... w^2 p^{-1/2} (fs_min = t^2 / 24.739...^2); evaluated as
    // (w_ab w_cd)^4 < fs_min^2 p q
    double fs_min;
    // per-quartet screens inside the team evaluator's quartet mode (lists
    // the routing kernel does not build: quartets with an f shell): the
    // density test with the shell-block maxima sblk (nullptr: off) and the
    // decay indicator sdecay (nullptr: off), the rules of the routing kernel
    const double* sblk;
    const int* spgA;
    const int* spgB;
    const double* sQ;
    int sng;
    int smode;          // 0 J and K, 1 J only, 2 K only
    double stau;
    const double* sdecay;
    double sdecay_thr;
    // team evaluator classes whose HRR and digestion work in global memory
    // (TeamShape::GWORK): one scratch slot per workgroup, gridDim.x <= slots
    double* gwork;
};

// Prefactor of the FASTSCHWARZ primitive-pair estimate, sqrt(2) pi^{5/2}
// (integrals/two_electron/rys/rys_engine.ipp:71-72).
constexpr double kFastSchwarzPrefactor = 24.739429451193148050228357586821944304008810699990039583;

struct StoreArgs {
    PairData pd;
    DevTables tb;
    const int2* quartets;   // {P, Q}
    const long long* off;   // output offset of each quartet's full block (nullptr: i * stride)
    long long stride;
    int n;
    double* out;
    double* gwork;          // team classes with global working sets: one slot per workgroup
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Lanes per ket of a task with nq kets: tasks with at most 32 kets give
// each ket ks lanes' worth of room (ks = nq rounded up to a power of two)
// and the 64 / ks lanes of a ket split its bra primitive pairs by stride.
__device__ __forceinline__ int ket_stride(int nq) {
    int ks = 1;
    while (ks < nq && ks < 64) ks <<= 1;
    return ks;
}

// One work item of a wave.  Task = {P, q_begin, q_end, bra primitive slice
// ib0 | ib1 << 16} (range mode) or a list-mode task (see JKArgs).  Lane l
// takes ket l % ks of the task (ks = ket_stride(kets)) and share l / ks of
// the 64 / ks shares of its primitive quartets (see quartet).
template <typename R, int LA, int LB, int LC, int LD, int AC, int AN, int CC, int CN, bool FS = false>
__device__ __forceinline__ void eri_jk_task(const JKArgs& a, const int wave, const double* sb) {
    using S = Shape<LA, LB, LC, LD, AC, AN, CC, CN>;
    const int lane = threadIdx.x & 63;
    const int task = a.qlist && a.split > 0 ? wave / a.split : wave;
    if (task >= a.ntasks) return;  // whole wave leaves together
    const int4 t = a.tasks[task];
    // the task is wave-uniform: keep it in scalar registers
    const int P = __builtin_amdgcn_readfirstlane(t.x);
    bool active;
    int Q, ib0, ib1, nq, ks;
    if (a.qlist) {
        // list mode: up to 64 kets at qlist[offset ..]; bra primitive slice
        // (wave % split) of split near-equal slices, or (split == 0) the
        // task's own slice [ib0, ib1) = (w & 0xffff, w >> 16)
        const long long off = static_cast<long long>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(t.y))) |
                              (static_cast<long long>(__builtin_amdgcn_readfirstlane(t.z) >> 8) << 32);
        nq = __builtin_amdgcn_readfirstlane(t.z) & 255;
        if (nq <= 0) return;
        const int pbn = a.pd.pn[P];
        if (a.split > 0) {
            const int ns = min(a.split, pbn);
            const int slice = wave - task * a.split;
            if (slice >= ns) return;
            const int step = (pbn + ns - 1) / ns;
            ib0 = slice * step;
            ib1 = min(pbn, ib0 + step);
        } else {
            const int tw = __builtin_amdgcn_readfirstlane(t.w);
            ib0 = tw & 0xffff;
            ib1 = min(pbn, (tw >> 16) & 0xffff);
        }
        if (ib0 >= ib1) return;
        ks = ket_stride(nq);
        const int kq = lane & (ks - 1);
        active = kq < nq;
        Q = a.qlist[off + (active ? kq : nq - 1)];
    } else {
        const int qb = __builtin_amdgcn_readfirstlane(t.y);
        int qe = __builtin_amdgcn_readfirstlane(t.z);
        const int tw = __builtin_amdgcn_readfirstlane(t.w);
        if (a.cut) qe = min(qe, a.cut[6 * P]);  // density screen (per bra pair, this ket class)
        if (qe <= qb) return;                     // whole wave leaves together
        nq = min(qe - qb, 64);
        ks = ket_stride(nq);
        const int qraw = qb + (lane & (ks - 1));
        active = qraw < qe;
        Q = active ? qraw : qe - 1;
        ib0 = tw & 0xffff;
        ib1 = (tw >> 16) & 0xffff;
    }
    const int ibs = 64 / ks;    // lanes per ket
    const int sub = lane / ks;  // this lane's share of the ket's primitive quartets
    constexpr bool flat = flat_prims(LA, LB, LC, LD);
    const int nshare = flat ? a.pd.pn[Q] * (ib1 - ib0) : ib1 - ib0;  // units the lanes of a ket split
    active = active && sub < nshare;
    if (a.counter && AC <= 0 && CC <= 0) {  // the class's first output chunk counts
        int npq = active ? (nshare - sub + ibs - 1) / ibs * (flat ? 1 : a.pd.pn[Q]) : 0;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) npq += __shfl_xor(npq, m, 64);
        if (lane == 0) {
            constexpr int base = std::is_same<R, float>::value ? 2 : 0;
            unsigned long long* cnt = a.counter + 4 * (wave % kCounterSlots);
            if (ib0 == 0) {
                atomicAdd(cnt, static_cast<unsigned long long>(nq));
                if (base) atomicAdd(cnt + base, static_cast<unsigned long long>(nq));
            }
            atomicAdd(cnt + 1, static_cast<unsigned long long>(npq));
            if (base) atomicAdd(cnt + base + 1, static_cast<unsigned long long>(npq));
        }
    }

    R acc[S::NOUT];
    quartet<R, LA, LB, LC, LD, AC, AN, CC, CN, FS>(a.tb, a.pd, P, Q, ib0, active ? ib1 : ib0, sub, ibs, true, acc,
                                                   a.fs_min, sb);
    if (ibs > 1) {
        // the ibs lanes of a ket (lane bits above ks) hold parts of one block:
        // sum them into the first, which alone digests
#pragma unroll
        for (int o = 0; o < S::NOUT; ++o) {
            R v = acc[o];
            for (int m = ks; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
            acc[o] = v;
        }
        active = (lane < ks) && (lane < nq);
    }

    if (a.cw) {
        // canonical-stream checksum: the weights fold the spherical projection
        // and the unique-index restriction of each pair into one factor
        constexpr int XB = LA == 0 ? 0 : (LA == 1 ? (LB == 0 ? 1 : 2) : (LB == 0 ? 3 : (LB == 1 ? 4 : 5)));
        constexpr int YK = LC == 0 ? 0 : (LC == 1 ? (LD == 0 ? 1 : 2) : (LD == 0 ? 3 : (LD == 1 ? 4 : 5)));
        const double* wp = a.cw + (2 * XB + (a.pd.diag[P] ? 1 : 0)) * 36;
        const double* wq = a.cw + (2 * YK + (a.pd.diag[Q] ? 1 : 0)) * 36;
        double v = 0.0;
        int o = 0;
#pragma unroll
        for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
            for (int b = 0; b < S::NB; ++b)
#pragma unroll
                for (int ci = 0; ci < S::NCL; ++ci)
#pragma unroll
                    for (int d = 0; d < S::ND; ++d, ++o)
                        v = fma(double(acc[o]) * wp[S::acomp(ai) * S::NB + b], wq[S::ccomp(ci) * S::ND + d], v);
        if (!active || P == Q) v = 0.0;
        v = wave_sum(v);
        if (lane == 0) unsafeAtomicAdd(a.chk, v);
        return;
    }

    // degeneracy scale: 1/2 per diagonal pair, 1/2 if P == Q
    double s = 1.0;
    if (a.pd.diag[P]) s *= 0.5;
    if (a.pd.diag[Q]) s *= 0.5;
    if (P == Q) s *= 0.5;
    if (!active) s = 0.0;

    long M = a.M;
    const double* __restrict__ Dm = a.D;
    double* __restrict__ Jm = a.J;
    double* __restrict__ Km = a.K;
    if (a.pmol) {  // both pairs of a task belong to P's molecule
        const int mol = __builtin_amdgcn_readfirstlane(a.pmol[P]);
        const long long base = a.mbase[mol];
        M = a.mM[mol];
        Dm += base;
        Jm += base;
        Km += base;
    }
    const int A0 = a.pd.ao[2 * P], B0 = a.pd.ao[2 * P + 1];
    const int C0 = a.pd.ao[2 * Q], D0 = a.pd.ao[2 * Q + 1];

    if (a.do_j) {
        // ket side: Jacc[c][d] += 2 s sum_ab (ab|cd) D[a][b]
        // bra side: Jacc[a][b] += 2 s sum_cd (ab|cd) D[c][d], reduced over the wave
        double dab[S::NAL][S::NB];
#pragma unroll
        for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
            for (int b = 0; b < S::NB; ++b) dab[ai][b] = Dm[(A0 + S::acomp(ai)) * M + B0 + b];
        double jab[S::NAL][S::NB];
#pragma unroll
        for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
            for (int b = 0; b < S::NB; ++b) jab[ai][b] = 0.0;
#pragma unroll
        for (int ci = 0; ci < S::NCL; ++ci)
#pragma unroll
            for (int d = 0; d < S::ND; ++d) {
                const long row = (C0 + S::ccomp(ci)) * M + D0 + d;
                const double dcd = Dm[row];
                double jcd = 0.0;
#pragma unroll
                for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
                    for (int b = 0; b < S::NB; ++b) {
                        const double v = double(acc[((ai * S::NB + b) * S::NCL + ci) * S::ND + d]);
                        jcd = fma(v, dab[ai][b], jcd);
                        jab[ai][b] = fma(v, dcd, jab[ai][b]);
                    }
                if (active) unsafeAtomicAdd(Jm + row, 2.0 * s * jcd);
            }
#pragma unroll
        for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
            for (int b = 0; b < S::NB; ++b) {
                const double v = wave_sum(2.0 * s * jab[ai][b]);
                if (lane == 0) unsafeAtomicAdd(Jm + (A0 + S::acomp(ai)) * M + B0 + b, v);
            }
    }
    if (a.do_k && active) {
        // Kacc[a][c] += s sum_{b,d} v D[b][d]; Kacc[b][c] += s sum_{a,d} v D[a][d]
        // Kacc[a][d] += s sum_{b,c} v D[b][c]; Kacc[b][d] += s sum_{a,c} v D[a][c]
#pragma unroll
        for (int ai = 0; ai < S::NAL; ++ai) {
            const int ga = A0 + S::acomp(ai);
#pragma unroll
            for (int ci = 0; ci < S::NCL; ++ci) {
                const int gc = C0 + S::ccomp(ci);
                double kac = 0.0;
#pragma unroll
                for (int b = 0; b < S::NB; ++b)
#pragma unroll
                    for (int d = 0; d < S::ND; ++d)
                        kac = fma(double(acc[((ai * S::NB + b) * S::NCL + ci) * S::ND + d]), Dm[(B0 + b) * M + D0 + d], kac);
                unsafeAtomicAdd(Km + ga * M + gc, s * kac);
            }
#pragma unroll
            for (int d = 0; d < S::ND; ++d) {
                const int gd = D0 + d;
                double kad = 0.0;
#pragma unroll
                for (int b = 0; b < S::NB; ++b)
#pragma unroll
                    for (int ci = 0; ci < S::NCL; ++ci)
                        kad = fma(double(acc[((ai * S::NB + b) * S::NCL + ci) * S::ND + d]),
                                  Dm[(B0 + b) * M + C0 + S::ccomp(ci)], kad);
                unsafeAtomicAdd(Km + ga * M + gd, s * kad);
            }
        }
#pragma unroll
        for (int b = 0; b < S::NB; ++b) {
            const int gb = B0 + b;
#pragma unroll
            for (int ci = 0; ci < S::NCL; ++ci) {
                const int gc = C0 + S::ccomp(ci);
                double kbc = 0.0;
#pragma unroll
                for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
                    for (int d = 0; d < S::ND; ++d)
                        kbc = fma(double(acc[((ai * S::NB + b) * S::NCL + ci) * S::ND + d]),
                                  Dm[(A0 + S::acomp(ai)) * M + D0 + d], kbc);
                unsafeAtomicAdd(Km + gb * M + gc, s * kbc);
            }
#pragma unroll
            for (int d = 0; d < S::ND; ++d) {
                const int gd = D0 + d;
                double kbd = 0.0;
#pragma unroll
                for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
                    for (int ci = 0; ci < S::NCL; ++ci)
                        kbd = fma(double(acc[((ai * S::NB + b) * S::NCL + ci) * S::ND + d]),
                                  Dm[(A0 + S::acomp(ai)) * M + C0 + S::ccomp(ci)], kbd);
                unsafeAtomicAdd(Km + gb * M + gd, s * kbd);
            }
        }
    }
}

// Grid entry: one work item per wave, or a persistent loop over work items.
// FS: the instance with the primitive FASTSCHWARZ screen (a kernel of its
// own, so the unscreened kernels keep their registers and occupancy).
template <typename R, int LA, int LB, int LC, int LD, int AC, int AN, int CC, int CN, bool FS = false>
__device__ __forceinline__ void eri_jk_body(const JKArgs& a, const double* sb) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    // persistent: the grid is whole blocks of four waves, so a cap that is
    // not a multiple of four launches waves beyond the stride; they leave
    // (their items belong to the waves below the stride)
    if (a.stride > 0 && wave >= a.stride) return;
    const int end = a.stride > 0 ? a.nwork : wave + 1;
    const int step = a.stride > 0 ? a.stride : 1;
    for (int w = wave; w < end; w += step) eri_jk_task<R, LA, LB, LC, LD, AC, AN, CC, CN, FS>(a, w, sb);
}

// Writes the lane's slice of the unit-diagonal Cartesian block of (P|Q) into
// out[off + ((a*NB + b)*NC + c)*ND + d] (bra/ket shells in pair orientation).
template <int LA, int LB, int LC, int LD, int AC, int AN, int CC, int CN>
__device__ __forceinline__ void eri_store_body(const StoreArgs& a) {
    using S = Shape<LA, LB, LC, LD, AC, AN, CC, CN>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int2 pq = a.quartets[i];
    double acc[S::NOUT];
    quartet<double, LA, LB, LC, LD, AC, AN, CC, CN>(a.tb, a.pd, pq.x, pq.y, 0, a.pd.pn[pq.x], 0, 1, false, acc);
    double* out = a.out + (a.off ? a.off[i] : static_cast<long long>(i) * a.stride);
    int o = 0;
#pragma unroll
    for (int ai = 0; ai < S::NAL; ++ai)
#pragma unroll
        for (int b = 0; b < S::NB; ++b)
#pragma unroll
            for (int ci = 0; ci < S::NCL; ++ci)
#pragma unroll
                for (int d = 0; d < S::ND; ++d, ++o)
                    out[((S::acomp(ai) * S::NB + b) * S::NC + S::ccomp(ci)) * S::ND + d] = acc[o];
}

// ------------------------------------------------------------ dispatch ---
using JKKernel = void (*)(JKArgs);
using StoreKernel = void (*)(StoreArgs);

// The output chunks of a class are compiled into groups of a few chunks per
// kernel (blockIdx.y selects the chunk inside the group): one kernel holding
// every chunk would be register-allocated for all of them at once.
constexpr int kMaxChunkGroups = 36;
struct JKGroup {
    JKKernel jk = nullptr;    // FP64 recurrences
    JKKernel jk32 = nullptr;  // FP32 recurrences (mixed-precision route)
    JKKernel jkfs = nullptr;  // the same two with the primitive FASTSCHWARZ screen
    JKKernel jk32fs = nullptr;
    int c0 = 0, n = 0;        // chunks [c0, c0 + n)
};
// Class pairs with an f shell run the team evaluator only: its J/K kernel
// and its block store (Schwarz diagonal and audit blocks).
struct HighKernels {
    JKKernel team = nullptr;
    StoreKernel store = nullptr;
    int nout = 0;
    int threads = 64;  // threads per quartet (workgroup size): 64 to 384
    long long gwork = 0;  // doubles of global scratch per workgroup (0: LDS only)
};
// pair classes: ss ps pp ds dp dd (the lane and team evaluators), then
// fs fp fd ff, gs gp gd gf gg (team evaluator only)
constexpr int kNumLowClasses = 6;
constexpr int kNumClasses = 15;
constexpr int kMaxShellL = 4;

struct ClassKernels {
    int nchunk = 0;
    int ngroup = 0;
    JKGroup grp[kMaxChunkGroups];
    StoreKernel store = nullptr;
    JKKernel team = nullptr;  // one wave per shell quartet, FP64 (classes with L_total >= 4)
    int nout = 0;             // Cartesian values per shell quartet
};

// class index: 0 ss, 1 ps, 2 pp, 3 ds, 4 dp, 5 dd (pair orientation: higher L first)
#define TLNO_AMD_REG(tag) void register_kernels_bra##tag(ClassKernels (&tab)[6][6]);
TLNO_AMD_REG(0)
TLNO_AMD_REG(1)
TLNO_AMD_REG(2)
TLNO_AMD_REG(3)
TLNO_AMD_REG(4)
TLNO_AMD_REG(4b)
TLNO_AMD_REG(5)
TLNO_AMD_REG(5b)
TLNO_AMD_REG(5c)
#undef TLNO_AMD_REG
// pair classes with an f or g shell: 6 fs, 7 fp, 8 fd, 9 ff, 10 gs, 11 gp,
// 12 gd, 13 gf, 14 gg (team evaluator only)
#define TLNO_AMD_REG_HI(tag) void register_kernels_##tag(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]);
TLNO_AMD_REG_HI(hi6)
TLNO_AMD_REG_HI(hi7)
TLNO_AMD_REG_HI(hi8a)
TLNO_AMD_REG_HI(hi8b)
TLNO_AMD_REG_HI(hi9a)
TLNO_AMD_REG_HI(hi9b)
TLNO_AMD_REG_HI(hi9c)
TLNO_AMD_REG_HI(hi10)
TLNO_AMD_REG_HI(hi11a)
TLNO_AMD_REG_HI(hi11b)
TLNO_AMD_REG_HI(hi12a)
TLNO_AMD_REG_HI(hi12b)
TLNO_AMD_REG_HI(hi12c)
TLNO_AMD_REG_HI(hi13a)
TLNO_AMD_REG_HI(hi13b)
TLNO_AMD_REG_HI(hi13c)
TLNO_AMD_REG_HI(hi13d)
TLNO_AMD_REG_HI(hi14a)
TLNO_AMD_REG_HI(hi14b)
TLNO_AMD_REG_HI(hi14c)
TLNO_AMD_REG_HI(hi14d)
TLNO_AMD_REG_HI(hi14e)
TLNO_AMD_REG_HI(hi14f)
#undef TLNO_AMD_REG_HI

}  // namespace tlno_amd
