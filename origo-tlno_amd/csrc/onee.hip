// One-electron S, T, V on the GPU.
//
// One wave per kept shell pair (bra-side shell has the higher L), one kernel
// instance per pair class.  The lanes split the primitive pairs for S and T
// and the (primitive pair, nucleus) combinations for V, and the partial
// blocks are summed across the wave with cross-lane shuffles:
//   * per primitive pair: Obara-Saika 1D overlap table S(i, j)
//     (i <= la + 1, j <= lb + 1) and the kinetic term in the symmetric form
//     1/2 <d/dx a|d/dx b>,
//     T_x = 1/2 [ij S(i-1,j-1) - 2ib S(i-1,j+1) - 2ja S(i+1,j-1) + 4ab S(i+1,j+1)]
//     (the reference's one-sided form, integrals/one_electron/
//     onee_integrals.cpp:283-448, cancels to a relative 1/b for a tight ket
//     against a diffuse bra and loses up to three digits at exponent ratios of
//     5000/0.08);
//   * per primitive pair and nucleus C: nuclear attraction by Rys quadrature
//     over the same device roots as the ERI engine, T = p |P - C|^2, per root
//     u = t^2 the 1D recursion G(i+1) = (PA - u PC) G(i) + i (1 - u)/(2p) G(i-1)
//     and the transfer G(i, j+1) = G(i+1, j) + AB G(i, j),
//     V = -Z (2 pi/p) K_AB sum_r w_r Gx Gy Gz.  (The reference evaluates V by
//     McMurchie-Davidson Hermite expansion, onee_integrals.cpp:476-918; both
//     are exact quadratures of the same integral.)
// Screening follows the reference (onee_matrices.cpp:100-210,
// onee_integrals.cpp:250-256, 541-556): the host drops group pairs below the
// decay estimate and marks pairs whose overlap bound is below it as S/T only;
// the kernel skips the V contribution of a primitive pair whose monopole
// estimate (2 pi/p) K_AB max|c_a| max|c_b| is below eps / sum|Z|.
// Blocks are written into internal-order matrices and mapped to the caller's
// Cartesian or spherical basis by the symmetrizing finalize kernels of the
// J/K path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "eri_device.hpp"
#include "onee.hpp"

namespace tlno_amd {

struct OneeArgs {
    PairData pd;
    DevTables tb;
    const int2* list;    // kept pairs {P, V wanted} of this class
    int nlist;
    const double* nucR;  // [nnuc][3]
    const double* nucZ;
    int nnuc;
    double vthr;         // nuclear primitive-pair screen threshold (eps / sum|Z|; 0 = off)
    double namax;        // max angular normalization of the bra times the ket shell's
    int M;
    double* S;
    double* T;
    double* V;
    const double* expA;  // per primitive pair: alpha, beta  [nprim][2]
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void k_onee(OneeArgs a) {
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr CartT<LA> ca = make_cart<LA>();
    constexpr CartT<LB> cb = make_cart<LB>();
    constexpr int NR = (LA + LB) / 2 + 1;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= a.nlist) return;
    const int2 ent = a.list[w];
    const int P = __builtin_amdgcn_readfirstlane(ent.x);
    const bool do_v = __builtin_amdgcn_readfirstlane(ent.y) != 0;
    double s[NA * NB], t[NA * NB], v[NA * NB];
#pragma unroll
    for (int o = 0; o < NA * NB; ++o) s[o] = t[o] = v[o] = 0.0;
    const double A[3] = {a.pd.A[3 * P], a.pd.A[3 * P + 1], a.pd.A[3 * P + 2]};
    const double AB[3] = {a.pd.AB[3 * P], a.pd.AB[3 * P + 1], a.pd.AB[3 * P + 2]};
    const int p0 = a.pd.poff[P], pn = a.pd.pn[P];
    // overlap and kinetic energy: one primitive pair per lane
    for (int k = lane; k < pn; k += 64) {
        const double4 Pp = a.pd.pP[p0 + k];
        const double p = Pp.w;
        const double al = a.expA[2 * (p0 + k)], be = a.expA[2 * (p0 + k) + 1];
        const double wK = a.pd.pw[p0 + k] * p;  // K_AB c_a c_b
        const double Pv[3] = {Pp.x, Pp.y, Pp.z};
        const double i2p = 0.5 / p;
        double S1[3][LA + 2][LB + 2];
        const double sp = sqrt(3.14159265358979323846 / p);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double PA = Pv[x] - A[x];
            const double PB = PA + AB[x];  // P - B = P - A + (A - B)
            // two-index Obara-Saika recursion (no transfer step: the transfer
            // S(i, j+1) = S(i+1, j) + AB S(i, j) cancels badly for a tight
            // function far from a diffuse one)
            //   S(i+1, j) = PA S(i,j) + (i S(i-1,j) + j S(i,j-1)) / (2p)
            //   S(i, j+1) = PB S(i,j) + (i S(i-1,j) + j S(i,j-1)) / (2p)
            double (&W)[LA + 2][LB + 2] = S1[x];
#pragma unroll
            for (int j = 0; j <= LB + 1; ++j)
#pragma unroll
                for (int i = 0; i <= LA + 1; ++i) {
                    double val;
                    if (i == 0 && j == 0) {
                        val = sp;  // exp(-mu AB^2) is folded into wK
                    } else if (i > 0) {
                        val = PA * W[i - 1][j];
                        if (i > 1) val += (i - 1) * i2p * W[i > 1 ? i - 2 : 0][j];
                        if (j > 0) val += j * i2p * W[i - 1][j > 0 ? j - 1 : 0];
                    } else {
                        val = PB * W[0][j - 1];
                        if (j > 1) val += (j - 1) * i2p * W[0][j > 1 ? j - 2 : 0];
                    }
                    W[i][j] = val;
                }
        }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                double sx[3], tx[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int i = ca.e[x][ia], j = cb.e[x][ib];
                    sx[x] = S1[x][i][j];
                    // 1/2 <d/dx a | d/dx b>: no cancellation between a tight
                    // and a diffuse exponent
                    double tv = 4.0 * al * be * S1[x][i + 1][j + 1];
                    if (i > 0) tv -= 2.0 * i * be * S1[x][i > 0 ? i - 1 : 0][j + 1];
                    if (j > 0) tv -= 2.0 * j * al * S1[x][i + 1][j > 0 ? j - 1 : 0];
                    if (i > 0 && j > 0) tv += double(i * j) * S1[x][i > 0 ? i - 1 : 0][j > 0 ? j - 1 : 0];
                    tx[x] = 0.5 * tv;
                }
                s[ia * NB + ib] += wK * sx[0] * sx[1] * sx[2];
                t[ia * NB + ib] += wK * (tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2] + sx[0] * sx[1] * tx[2]);
            }
    }
    // nuclear attraction: one (primitive pair, nucleus) combination per lane,
    // nuclei fastest (the lanes of a wave share primitive pairs)
    const int nn = a.nnuc;
    const int nitem = do_v ? pn * nn : 0;
    for (int it = lane; it < nitem; it += 64) {
        const int k = it / nn, c = it - k * nn;
        const double4 Pp = a.pd.pP[p0 + k];
        const double p = Pp.w;
        const double wK = a.pd.pw[p0 + k] * p;
        const double twopi_p = 2.0 * 3.14159265358979323846 / p;
        if (twopi_p * fabs(wK) * a.namax < a.vthr) continue;
        const double Pv[3] = {Pp.x, Pp.y, Pp.z};
        const double i2p = 0.5 / p;
        const double PC[3] = {Pv[0] - a.nucR[3 * c], Pv[1] - a.nucR[3 * c + 1], Pv[2] - a.nucR[3 * c + 2]};
        const double Targ = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        double rt[NR], wt[NR];
        rys_roots<NR>(a.tb, Targ, rt, wt);
        const double pre = -a.nucZ[c] * twopi_p * wK;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double u = rt[r];
            const double B = (1.0 - u) * i2p;
            double I[3][LA + 1][LB + 1];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const double C0 = (Pv[x] - A[x]) - u * PC[x];
                double G[LA + LB + 1][LB + 1];
                G[0][0] = (x == 2) ? pre * wt[r] : 1.0;
#pragma unroll
                for (int i = 0; i < LA + LB; ++i) G[i + 1][0] = C0 * G[i][0] + (i > 0 ? i * B * G[i - 1][0] : 0.0);
#pragma unroll
                for (int j = 0; j < LB; ++j)
#pragma unroll
                    for (int i = 0; i + j + 1 <= LA + LB; ++i) G[i][j + 1] = G[i + 1][j] + AB[x] * G[i][j];
#pragma unroll
                for (int i = 0; i <= LA; ++i)
#pragma unroll
                    for (int j = 0; j <= LB; ++j) I[x][i][j] = G[i][j];
            }
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
                    v[ia * NB + ib] += I[0][ca.e[0][ia]][cb.e[0][ib]] * I[1][ca.e[1][ia]][cb.e[1][ib]] *
                                       I[2][ca.e[2][ia]][cb.e[2][ib]];
        }
    }
    const double half = a.pd.diag[P] ? 0.5 : 1.0;
    const long M = a.M;
    const int a0 = a.pd.ao[2 * P], b0 = a.pd.ao[2 * P + 1];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const int o = ia * NB + ib;
            const double so = wave_sum(s[o]), to = wave_sum(t[o]), vo = wave_sum(v[o]);
            if (lane == (o & 63)) {
                const double f = half * ca.nang[ia] * cb.nang[ib];
                const long idx = (a0 + ia) * M + b0 + ib;
                a.S[idx] = f * so;
                a.T[idx] = f * to;
                a.V[idx] = f * vo;
            }
        }
}

// max over the Cartesian components of a complete shell of
// 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)
double onee_namax(int L) {
    double best = 0.0;
    for (int x = L; x >= 0; --x)
        for (int y = L - x; y >= 0; --y) {
            const int z = L - x - y;
            best = std::max(best, 1.0 / std::sqrt(double(dfact2m1(x) * dfact2m1(y) * dfact2m1(z))));
        }
    return best;
}

void launch_onee(const PairData& pd, const DevTables& tb, const int* dL, const double* dexp, const double* dR,
                 const double* dZ, int nnuc, const int2* dlist, const int* class_begin, double vthr, int M,
                 double* S, double* T, double* V, hipStream_t st) {
    (void)dL;
    static const int LAB[kNumClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}, {3, 0}, {3, 1},
                                            {3, 2}, {3, 3}, {4, 0}, {4, 1}, {4, 2}, {4, 3}, {4, 4}};
    for (int c = 0; c < kNumClasses; ++c) {
        const int n = class_begin[c + 1] - class_begin[c];
        if (n <= 0) continue;
        OneeArgs a{};
        a.pd = pd;
        a.tb = tb;
        a.list = dlist + class_begin[c];
        a.nlist = n;
        a.nucR = dR;
        a.nucZ = dZ;
        a.nnuc = nnuc;
        a.vthr = vthr;
        a.namax = onee_namax(LAB[c][0]) * onee_namax(LAB[c][1]);
        a.M = M;
        a.S = S;
        a.T = T;
        a.V = V;
        a.expA = dexp;
        const unsigned grid = static_cast<unsigned>((n + 3) / 4);
        switch (c) {
            case 0: k_onee<0, 0><<<grid, 256, 0, st>>>(a); break;
            case 1: k_onee<1, 0><<<grid, 256, 0, st>>>(a); break;
            case 2: k_onee<1, 1><<<grid, 256, 0, st>>>(a); break;
            case 3: k_onee<2, 0><<<grid, 256, 0, st>>>(a); break;
            case 4: k_onee<2, 1><<<grid, 256, 0, st>>>(a); break;
            case 5: k_onee<2, 2><<<grid, 256, 0, st>>>(a); break;
            case 6: k_onee<3, 0><<<grid, 256, 0, st>>>(a); break;
            case 7: k_onee<3, 1><<<grid, 256, 0, st>>>(a); break;
            case 8: k_onee<3, 2><<<grid, 256, 0, st>>>(a); break;
            case 9: k_onee<3, 3><<<grid, 256, 0, st>>>(a); break;
            case 10: k_onee<4, 0><<<grid, 256, 0, st>>>(a); break;
            case 11: k_onee<4, 1><<<grid, 256, 0, st>>>(a); break;
            case 12: k_onee<4, 2><<<grid, 256, 0, st>>>(a); break;
            case 13: k_onee<4, 3><<<grid, 256, 0, st>>>(a); break;
            default: k_onee<4, 4><<<grid, 256, 0, st>>>(a); break;
        }
    }
}

}  // namespace tlno_amd
