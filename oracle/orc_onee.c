/*
 * ORACLE (test infrastructure only): one-electron S, T, V.
 *
 * Follows integrals/one_electron/onee_integrals.cpp (reference):
 *  - overlap from 1D Gaussian-product tables (32-51, 283-448);
 *  - kinetic T = T_x S_y S_z + S_x T_y S_z + S_x S_y T_z (283-448).  The
 *    reference applies -1/2 d^2/dx^2 to the ket,
 *    T_x = b(2j+1) S(i,j) - 2 b^2 S(i,j+2) - j(j-1)/2 S(i,j-2), whose terms
 *    cancel to a relative 1/b when a tight ket meets a diffuse bra; the
 *    oracle uses the equal symmetric form 1/2 <d/dx a|d/dx b>,
 *    T_x = 1/2 [ij S(i-1,j-1) - 2ib S(i-1,j+1) - 2ja S(i+1,j-1) + 4ab S(i+1,j+1)];
 *  - nuclear attraction by McMurchie-Davidson: Hermite E^{ij}_t per axis,
 *    R^{(n)}_{tuv} recursion from R^{(n)}_{000} = (-2p)^n F_n(p |PC|^2),
 *    V = -(2 pi / p) sum_C Z_C sum_tuv E_t E_u E_v R_tuv (53-157, 476-918).
 * Coefficients are the stored per-component coefficients (radial contraction
 * times N_A), output in the Cartesian component order or transformed to the
 * spherical basis with the same projector as the two-electron oracle.
 */
#include "oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793238462643383279502884

/* E[i][j][t] for i <= imax, j <= jmax (t <= i + j) */
static void hermite_E(int imax, int jmax, double p, double PA, double PB, double Kx,
                      double E[8][8][16]) {
    memset(E, 0, sizeof(double) * 8 * 8 * 16);
    E[0][0][0] = Kx;
    const double h = 0.5 / p;
    for (int i = 0; i <= imax; ++i) {
        if (i > 0)
            for (int t = 0; t <= i; ++t) {
                double v = PA * E[i - 1][0][t];
                if (t > 0) v += h * E[i - 1][0][t - 1];
                if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
                E[i][0][t] = v;
            }
        for (int j = 1; j <= jmax; ++j)
            for (int t = 0; t <= i + j; ++t) {
                double v = PB * E[i][j - 1][t];
                if (t > 0) v += h * E[i][j - 1][t - 1];
                if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
                E[i][j][t] = v;
            }
    }
}

static void prim_block(const orc_group* A, const orc_group* B, int ia, int ib, int n_nuc,
                       const double* Z, const double* R, double* S, double* Tk, double* V) {
    const double al = A->exps[ia], be = B->exps[ib];
    const double p = al + be, mu = al * be / p;
    double P[3], PA[3], PB[3];
    double E[3][8][8][16];
    for (int k = 0; k < 3; ++k) {
        P[k] = A->center[k] - be / p * (A->center[k] - B->center[k]);
        PA[k] = P[k] - A->center[k];
        PB[k] = P[k] - B->center[k];
        const double X = A->center[k] - B->center[k];
        hermite_E(A->L + 1, B->L + 1, p, PA[k], PB[k], exp(-mu * X * X), E[k]);
    }
    const double sp = sqrt(PI_D / p);
    const int na = A->ncomp, nb = B->ncomp;
    /* overlap/kinetic */
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            double s1[3], t1[3];
            for (int k = 0; k < 3; ++k) {
                const int i = A->ang[a][k], j = B->ang[b][k];
                s1[k] = E[k][i][j][0] * sp;
                /* symmetric (gradient) form, well conditioned for a tight
                 * function against a diffuse one */
                double tv = 4.0 * al * be * E[k][i + 1][j + 1][0];
                if (i > 0) tv -= 2.0 * i * be * E[k][i - 1][j + 1][0];
                if (j > 0) tv -= 2.0 * j * al * E[k][i + 1][j - 1][0];
                if (i > 0 && j > 0) tv += (double)(i * j) * E[k][i - 1][j - 1][0];
                t1[k] = 0.5 * tv * sp;
            }
            S[a * nb + b] = s1[0] * s1[1] * s1[2];
            Tk[a * nb + b] = t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2];
        }
    /* nuclear */
    const int L = A->L + B->L;
    for (int x = 0; x < na * nb; ++x) V[x] = 0.0;
    for (int c = 0; c < n_nuc; ++c) {
        const double PC[3] = {P[0] - R[3 * c], P[1] - R[3 * c + 1], P[2] - R[3 * c + 2]};
        const double x = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        /* Rn[n][t][u][v] */
        static __thread double Rn[9][9][9][9];
        double Fn[9];
        orc_boys_all(L, x, Fn);
        double pw = 1.0;
        for (int n = 0; n <= L; ++n) {
            Rn[n][0][0][0] = pw * Fn[n];
            pw *= -2.0 * p;
        }
        for (int n = L - 1; n >= 0; --n) {
            const int lmax = L - n;
            for (int t = 0; t <= lmax; ++t)
                for (int u = 0; u + t <= lmax; ++u)
                    for (int v = 0; v + u + t <= lmax; ++v) {
                        if (t + u + v == 0) continue;
                        double val;
                        if (t > 0) {
                            val = PC[0] * Rn[n + 1][t - 1][u][v];
                            if (t > 1) val += (t - 1) * Rn[n + 1][t - 2][u][v];
                        } else if (u > 0) {
                            val = PC[1] * Rn[n + 1][t][u - 1][v];
                            if (u > 1) val += (u - 1) * Rn[n + 1][t][u - 2][v];
                        } else {
                            val = PC[2] * Rn[n + 1][t][u][v - 1];
                            if (v > 1) val += (v - 1) * Rn[n + 1][t][u][v - 2];
                        }
                        Rn[n][t][u][v] = val;
                    }
        }
        const double pre = -Z[c] * 2.0 * PI_D / p;
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) {
                const int* ta = A->ang[a];
                const int* tb = B->ang[b];
                double sum = 0.0;
                for (int t = 0; t <= ta[0] + tb[0]; ++t)
                    for (int u = 0; u <= ta[1] + tb[1]; ++u)
                        for (int v = 0; v <= ta[2] + tb[2]; ++v)
                            sum += E[0][ta[0]][tb[0]][t] * E[1][ta[1]][tb[1]][u] * E[2][ta[2]][tb[2]][v] *
                                   Rn[0][t][u][v];
                V[a * nb + b] += pre * sum;
            }
    }
}

int orc_onee(const orc_basis* b, int spherical, int n_nuc, const double* Z,
             const double* R, double* S, double* T, double* V) {
    const int n = spherical ? b->nsph : b->ncart;
    memset(S, 0, sizeof(double) * (size_t)n * n);
    memset(T, 0, sizeof(double) * (size_t)n * n);
    memset(V, 0, sizeof(double) * (size_t)n * n);
    for (int gi = 0; gi < b->ngroups; ++gi)
        for (int gj = 0; gj < b->ngroups; ++gj) {
            const orc_group* A = &b->groups[gi];
            const orc_group* B = &b->groups[gj];
            const int na = A->ncomp, nb = B->ncomp;
            double bs[225], bt[225], bv[225], ps[225], pt[225], pv[225];
            memset(bs, 0, sizeof bs);
            memset(bt, 0, sizeof bt);
            memset(bv, 0, sizeof bv);
            for (int i = 0; i < A->nprim; ++i)
                for (int j = 0; j < B->nprim; ++j) {
                    prim_block(A, B, i, j, n_nuc, Z, R, ps, pt, pv);
                    const double cc = A->c_radial[i] * B->c_radial[j];
                    for (int x = 0; x < na * nb; ++x) {
                        bs[x] += cc * ps[x];
                        bt[x] += cc * pt[x];
                        bv[x] += cc * pv[x];
                    }
                }
            for (int a = 0; a < na; ++a)
                for (int c = 0; c < nb; ++c) {
                    const double f = A->n_ang[a] * B->n_ang[c];
                    bs[a * nb + c] *= f;
                    bt[a * nb + c] *= f;
                    bv[a * nb + c] *= f;
                }
            if (!spherical) {
                for (int a = 0; a < na; ++a)
                    for (int c = 0; c < nb; ++c) {
                        const size_t o = (size_t)A->comp[a] * n + B->comp[c];
                        S[o] = bs[a * nb + c];
                        T[o] = bt[a * nb + c];
                        V[o] = bv[a * nb + c];
                    }
            } else {
                const int sa = 2 * A->L + 1, sb = 2 * B->L + 1;
                for (int s = 0; s < sa; ++s)
                    for (int t = 0; t < sb; ++t) {
                        double vs = 0, vt = 0, vv = 0;
                        for (int a = 0; a < na; ++a)
                            for (int c = 0; c < nb; ++c) {
                                const double w = A->c2s_comp[s * na + a] * B->c2s_comp[t * nb + c];
                                vs += w * bs[a * nb + c];
                                vt += w * bt[a * nb + c];
                                vv += w * bv[a * nb + c];
                            }
                        const size_t o = (size_t)(A->sph_base + s) * n + (B->sph_base + t);
                        S[o] = vs;
                        T[o] = vt;
                        V[o] = vv;
                    }
            }
        }
    return 0;
}
