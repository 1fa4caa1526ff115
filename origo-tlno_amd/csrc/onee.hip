// One-electron S, T, V on the GPU.
//
// One thread per shell pair (bra-side shell has the higher L).  Per primitive
// pair: Obara-Saika 1D overlap table S(i, j) (i <= la + 1, j <= lb + 1), the
// kinetic term in the symmetric form 1/2 <d/dx a|d/dx b>,
// T_x = 1/2 [ij S(i-1,j-1) - 2ib S(i-1,j+1) - 2ja S(i+1,j-1) + 4ab S(i+1,j+1)]
// (the reference's one-sided form, integrals/one_electron/onee_integrals.cpp:
// 283-448, cancels to a relative 1/b for a tight ket against a diffuse bra
// and loses up to three digits at exponent ratios of 5000/0.08), and the nuclear
// attraction by Rys quadrature over the same device roots as the ERI engine:
// for each nucleus C, T = p |P - C|^2, per root u = t^2 the 1D recursion
// G(i+1) = (PA - u PC) G(i) + i (1 - u)/(2p) G(i-1) and the transfer
// G(i, j+1) = G(i+1, j) + AB G(i, j), V = -Z (2 pi/p) K_AB sum_r w_r Gx Gy Gz.
// (The reference evaluates V by McMurchie-Davidson Hermite expansion,
// onee_integrals.cpp:476-918; both are exact quadratures of the same
// integral.)  Blocks are written into internal-order matrices and mapped to
// the caller's Cartesian or spherical basis by the symmetrizing finalize
// kernels of the J/K path.
#include <hip/hip_runtime.h>

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
    const int* pL;       // [npair][2] LA, LB
    const double* prim;  // [npair] not used
    const double* nucR;  // [nnuc][3]
    const double* nucZ;
    int nnuc;
    int npair;
    int M;
    double* S;
    double* T;
    double* V;
    const double* expA; // per primitive pair: alpha, beta  [nprim][2]
};

template <int LA, int LB>
__device__ void onee_pair(const OneeArgs& a, int P) {
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr CartT<LA> ca = make_cart<LA>();
    constexpr CartT<LB> cb = make_cart<LB>();
    constexpr int NR = (LA + LB) / 2 + 1;
    double s[NA * NB], t[NA * NB], v[NA * NB];
#pragma unroll
    for (int o = 0; o < NA * NB; ++o) s[o] = t[o] = v[o] = 0.0;
    const double Ax = a.pd.A[3 * P], Ay = a.pd.A[3 * P + 1], Az = a.pd.A[3 * P + 2];
    const double AB[3] = {a.pd.AB[3 * P], a.pd.AB[3 * P + 1], a.pd.AB[3 * P + 2]};
    const double A[3] = {Ax, Ay, Az};
    const int p0 = a.pd.poff[P], pn = a.pd.pn[P];
    for (int k = 0; k < pn; ++k) {
        const double4 Pp = a.pd.pP[p0 + k];
        const double p = Pp.w;
        const double al = a.expA[2 * (p0 + k)], be = a.expA[2 * (p0 + k) + 1];
        const double wK = a.pd.pw[p0 + k] * p;  // K_AB c_a c_b
        const double Pv[3] = {Pp.x, Pp.y, Pp.z};
        const double i2p = 0.5 / p;
        // overlap tables per axis: Sx[i][j], i <= LA + 1, j <= LB + 1
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
                    double v;
                    if (i == 0 && j == 0) {
                        v = sp;  // exp(-mu AB^2) is folded into wK
                    } else if (i > 0) {
                        v = PA * W[i - 1][j];
                        if (i > 1) v += (i - 1) * i2p * W[i > 1 ? i - 2 : 0][j];
                        if (j > 0) v += j * i2p * W[i - 1][j > 0 ? j - 1 : 0];
                    } else {
                        v = PB * W[0][j - 1];
                        if (j > 1) v += (j - 1) * i2p * W[0][j > 1 ? j - 2 : 0];
                    }
                    W[i][j] = v;
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
        (void)al;
        // nuclear attraction
        for (int c = 0; c < a.nnuc; ++c) {
            const double PC[3] = {Pv[0] - a.nucR[3 * c], Pv[1] - a.nucR[3 * c + 1], Pv[2] - a.nucR[3 * c + 2]};
            const double Targ = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
            double rt[NR], wt[NR];
            rys_roots<NR>(a.tb, Targ, rt, wt);
            const double pre = -a.nucZ[c] * 2.0 * 3.14159265358979323846 / p * wK;
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
                    for (int i = 0; i < LA + LB; ++i)
                        G[i + 1][0] = C0 * G[i][0] + (i > 0 ? i * B * G[i - 1][0] : 0.0);
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
    }
    const double half = a.pd.diag[P] ? 0.5 : 1.0;
    const long M = a.M;
    const int a0 = a.pd.ao[2 * P], b0 = a.pd.ao[2 * P + 1];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const double f = half * ca.nang[ia] * cb.nang[ib];
            const long o = (a0 + ia) * M + b0 + ib;
            a.S[o] = f * s[ia * NB + ib];
            a.T[o] = f * t[ia * NB + ib];
            a.V[o] = f * v[ia * NB + ib];
        }
}

__global__ __launch_bounds__(128) void k_onee(OneeArgs a) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= a.npair) return;
    const int LA = a.pL[2 * P], LB = a.pL[2 * P + 1];
    if (LA == 0) onee_pair<0, 0>(a, P);
    else if (LA == 1 && LB == 0) onee_pair<1, 0>(a, P);
    else if (LA == 1) onee_pair<1, 1>(a, P);
    else if (LB == 0) onee_pair<2, 0>(a, P);
    else if (LB == 1) onee_pair<2, 1>(a, P);
    else onee_pair<2, 2>(a, P);
}

void launch_onee(const PairData& pd, const DevTables& tb, const int* dL, const double* dexp, const double* dR,
                 const double* dZ, int nnuc, int npair, int M, double* S, double* T, double* V, hipStream_t st) {
    OneeArgs a{};
    a.pd = pd;
    a.tb = tb;
    a.pL = dL;
    a.nucR = dR;
    a.nucZ = dZ;
    a.nnuc = nnuc;
    a.npair = npair;
    a.M = M;
    a.S = S;
    a.T = T;
    a.V = V;
    a.expA = dexp;
    k_onee<<<(npair + 127) / 128, 128, 0, st>>>(a);
}

}  // namespace tlno_amd
