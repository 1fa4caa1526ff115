/*
 * ORACLE (test infrastructure only): Rys-quadrature ERIs, Schwarz bounds and
 * direct J/K digestion.
 *
 * Primitive-pair data (rys_engine.ipp:2243-2277): p = a + b,
 * P = (a A + b B)/p, K_AB = exp(-a b/p |AB|^2), radial weight c_a c_b.
 * Per primitive quartet (rys_engine.ipp:13540-13628):
 *   T = rho |PQ|^2, rho = pq/(p+q), prefactor 2 pi^{5/2} K_AB K_CD c.../(pq sqrt(p+q));
 *   per root t^2: B00 = t^2/(2(p+q)), B10 = (1 - q t^2/(p+q))/(2p),
 *   B01 = (1 - p t^2/(p+q))/(2q), C00 = (P-A) - q/(p+q) PQ t^2,
 *   D00 = (Q-C) + p/(p+q) PQ t^2, 2D VRR
 *   G[i+1,0] = C00 G[i,0] + i B10 G[i-1,0],
 *   G[i,k+1] = D00 G[i,k] + i B00 G[i-1,k] + k B01 G[i,k-1];
 *   ket HRR (c,d+1) = (c+1,d) + CD (c,d), bra HRR (a,b+1) = (a+1,b) + AB (a,b)
 *   (rys_engine.ipp:3155-3363), then the Cartesian product Ix Iy Iz summed
 *   over roots and primitives.  Output is the unit-diagonal Cartesian block:
 *   c_radial per primitive and N_A per component (rys_engine.ipp:394-398,
 *   2016-2061).
 *
 * J/K (fock/jk_direct.hpp:87; consumer_policies.ipp:80-157):
 *   J[a,b] = sum_{c,d} P[c,d] (ab|cd),  K[a,c] = sum_{b,d} P[b,d] (ab|cd).
 * The oracle computes each canonical shell quartet once and visits every
 * DISTINCT ordered shell quadruple of its 8-fold orbit explicitly, so no
 * degeneracy factor enters.  Schwarz screening (rys_engine.ipp:16164-16258,
 * native_quartets.ipp:571-574): a canonical quartet is skipped when
 * Q_ab * Q_cd < threshold with Q the max over the pair's AO pairs of
 * sqrt|(ab|ab)| in the output basis.
 */
#include "oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int orc_rys_table_init(int N);
void orc_rys_fast(int N, double T, double* roots, double* weights);

#define PI_D 3.141592653589793238462643383279502884

/* raw (pre-N_A) block in component order [a][b][c][d] */
/* fs2 > 0: primitive FASTSCHWARZ screen (integrals/two_electron/rys/
 * native_quartets.ipp:595-600, rys_engine.ipp:2243-2277): a primitive quartet
 * is left out when fs_ab fs_cd < fs2, fs = 24.739... w^2 p^{-1/2},
 * w = K c_a c_b / p (radial coefficients), applied here to every primitive
 * quartet on its own. */
#define ORC_FASTSCHWARZ_PREFACTOR 24.739429451193148050228357586821944304008810699990039583
static void eri_block(const orc_group* A, const orc_group* B, const orc_group* C,
                      const orc_group* D, int fast_roots, double fs2, double* out) {
    const int la = A->L, lb = B->L, lc = C->L, ld = D->L;
    const int LAB = la + lb, LCD = lc + ld;
    const int N = (LAB + LCD) / 2 + 1;
    const int na = A->ncomp, nb = B->ncomp, nc = C->ncomp, nd = D->ncomp;
    const int nout = na * nb * nc * nd;
    memset(out, 0, sizeof(double) * (size_t)nout);
    const double AB[3] = {A->center[0] - B->center[0], A->center[1] - B->center[1],
                          A->center[2] - B->center[2]};
    const double CD[3] = {C->center[0] - D->center[0], C->center[1] - D->center[1],
                          C->center[2] - D->center[2]};
    const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const double rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
    const double two_pi52 = 2.0 * pow(PI_D, 2.5);
    double G[9][9], H[9][9][5];       /* G[e][f], H[e][c][d] */
    static __thread double I[3][5][5][5][5]; /* I[axis][a][b][c][d] */
    double roots[16], wts[16];

    for (int i = 0; i < A->nprim; ++i)
        for (int j = 0; j < B->nprim; ++j) {
            const double al = A->exps[i], be = B->exps[j];
            const double p = al + be;
            const double Kab = exp(-al * be / p * rab2);
            const double cab = A->c_radial[i] * B->c_radial[j];
            double P[3];
            /* P = (a A + b B)/p, evaluated as A - (b/p)(A - B) (exact for one centre) */
            for (int k = 0; k < 3; ++k) P[k] = A->center[k] - be / p * AB[k];
            for (int k2 = 0; k2 < C->nprim; ++k2)
                for (int l = 0; l < D->nprim; ++l) {
                    const double ga = C->exps[k2], de = D->exps[l];
                    const double q = ga + de;
                    const double Kcd = exp(-ga * de / q * rcd2);
                    const double ccd = C->c_radial[k2] * D->c_radial[l];
                    double Q[3], PQ[3];
                    for (int k = 0; k < 3; ++k) {
                        Q[k] = C->center[k] - de / q * CD[k];
                        PQ[k] = P[k] - Q[k];
                    }
                    if (fs2 > 0.0) {
                        const double wab = Kab * cab / p, wcd = Kcd * ccd / q;
                        const double fab = ORC_FASTSCHWARZ_PREFACTOR * wab * wab * sqrt(1.0 / p);
                        const double fcd = ORC_FASTSCHWARZ_PREFACTOR * wcd * wcd * sqrt(1.0 / q);
                        if (fab * fcd < fs2) continue;
                    }
                    const double pq = p + q;
                    const double rho = p * q / pq;
                    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    const double pref = two_pi52 / (p * q * sqrt(pq)) * Kab * Kcd * cab * ccd;
                    if (fast_roots) orc_rys_fast(N, T, roots, wts);
                    else orc_rys(N, T, roots, wts);
                    for (int r = 0; r < N; ++r) {
                        const double t2 = roots[r];
                        const double B00 = 0.5 * t2 / pq;
                        const double B10 = (1.0 - q * t2 / pq) / (2.0 * p);
                        const double B01 = (1.0 - p * t2 / pq) / (2.0 * q);
                        for (int ax = 0; ax < 3; ++ax) {
                            const double C00 = (P[ax] - A->center[ax]) - q / pq * PQ[ax] * t2;
                            const double D00 = (Q[ax] - C->center[ax]) + p / pq * PQ[ax] * t2;
                            G[0][0] = 1.0;
                            for (int e = 0; e < LAB; ++e)
                                G[e + 1][0] = C00 * G[e][0] + (e > 0 ? e * B10 * G[e - 1][0] : 0.0);
                            for (int f = 0; f < LCD; ++f)
                                for (int e = 0; e <= LAB; ++e) {
                                    double v = D00 * G[e][f];
                                    if (e > 0) v += e * B00 * G[e - 1][f];
                                    if (f > 0) v += f * B01 * G[e][f - 1];
                                    G[e][f + 1] = v;
                                }
                            /* ket HRR: H[e][c][d] */
                            for (int e = 0; e <= LAB; ++e) {
                                for (int c = 0; c <= LCD; ++c) H[e][c][0] = G[e][c];
                                for (int d = 0; d < ld; ++d)
                                    for (int c = 0; c + d + 1 <= LCD; ++c)
                                        H[e][c][d + 1] = H[e][c + 1][d] + CD[ax] * H[e][c][d];
                            }
                            /* bra HRR: J[a][b] over rows e */
                            for (int c = 0; c <= lc; ++c)
                                for (int d = 0; d <= ld; ++d) {
                                    double Jb[9][5];
                                    for (int e = 0; e <= LAB; ++e) Jb[e][0] = H[e][c][d];
                                    for (int b = 0; b < lb; ++b)
                                        for (int a = 0; a + b + 1 <= LAB; ++a)
                                            Jb[a][b + 1] = Jb[a + 1][b] + AB[ax] * Jb[a][b];
                                    for (int a = 0; a <= la; ++a)
                                        for (int b = 0; b <= lb; ++b) I[ax][a][b][c][d] = Jb[a][b];
                                }
                        }
                        const double ww = pref * wts[r];
                        int o = 0;
                        for (int ia = 0; ia < na; ++ia) {
                            const int* ta = A->ang[ia];
                            for (int ib = 0; ib < nb; ++ib) {
                                const int* tb = B->ang[ib];
                                for (int ic = 0; ic < nc; ++ic) {
                                    const int* tc = C->ang[ic];
                                    for (int id = 0; id < nd; ++id, ++o) {
                                        const int* td = D->ang[id];
                                        out[o] += ww * I[0][ta[0]][tb[0]][tc[0]][td[0]] *
                                                  I[1][ta[1]][tb[1]][tc[1]][td[1]] *
                                                  I[2][ta[2]][tb[2]][tc[2]][td[2]];
                                    }
                                }
                            }
                        }
                    }
                }
        }
    int o = 0;
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < nc; ++ic)
                for (int id = 0; id < nd; ++id, ++o)
                    out[o] *= A->n_ang[ia] * B->n_ang[ib] * C->n_ang[ic] * D->n_ang[id];
}

int orc_eri_cart_block(const orc_basis* b, int ga, int gb, int gc, int gd, double* out) {
    const int ng = b->ngroups;
    if (ga < 0 || gb < 0 || gc < 0 || gd < 0 || ga >= ng || gb >= ng || gc >= ng || gd >= ng) {
        orc_set_error("oracle: group index out of range");
        return 1;
    }
    eri_block(&b->groups[ga], &b->groups[gb], &b->groups[gc], &b->groups[gd], 0, 0.0, out);
    return 0;
}

/* transform the Cartesian block to the spherical basis (tmp holds 2 x 50625) */
static void block_to_sph(const orc_group* G[4], const double* cart, double* sph, double* tmp) {
    int dims[4];
    for (int k = 0; k < 4; ++k) dims[k] = G[k]->ncomp;
    const double* src = cart;
    for (int k = 0; k < 4; ++k) {
        double* dst = (k == 3) ? sph : tmp + (size_t)(k % 2) * 50625;
        int outer = 1, inner = 1;
        for (int m = 0; m < k; ++m) outer *= dims[m];
        for (int m = k + 1; m < 4; ++m) inner *= dims[m];
        const double* T = G[k]->c2s_comp;
        const int ni = dims[k], no = 2 * G[k]->L + 1;
        for (int o = 0; o < outer; ++o)
            for (int s = 0; s < no; ++s)
                for (int in = 0; in < inner; ++in) {
                    double v = 0.0;
                    for (int c = 0; c < ni; ++c)
                        v += T[s * ni + c] * src[((size_t)o * ni + c) * inner + in];
                    dst[((size_t)o * no + s) * inner + in] = v;
                }
        dims[k] = no;
        src = dst;
    }
}

/* block of group quartet in the output basis; returns dims and AO indices */
static void out_block(const orc_basis* b, int spherical, int gi, int gj, int gk, int gl, double fs2,
                      double* cart, double* outb, double* tmp, int* dims, int* idx[4]) {
    const orc_group* G[4] = {&b->groups[gi], &b->groups[gj], &b->groups[gk], &b->groups[gl]};
    eri_block(G[0], G[1], G[2], G[3], 1, fs2, cart);
    for (int k = 0; k < 4; ++k) {
        if (spherical) {
            dims[k] = 2 * G[k]->L + 1;
            for (int s = 0; s < dims[k]; ++s) idx[k][s] = G[k]->sph_base + s;
        } else {
            dims[k] = G[k]->ncomp;
            for (int c = 0; c < dims[k]; ++c) idx[k][c] = G[k]->comp[c];
        }
    }
    if (spherical) block_to_sph(G, cart, outb, tmp);
    else memcpy(outb, cart, sizeof(double) * (size_t)(dims[0] * dims[1] * dims[2] * dims[3]));
}

static void init_tables(const orc_basis* b) {
    int Lmax = 0;
    for (int g = 0; g < b->ngroups; ++g)
        if (b->groups[g].L > Lmax) Lmax = b->groups[g].L;
    for (int N = 1; N <= 2 * Lmax + 1 && N <= 9; ++N) orc_rys_table_init(N);
}

int orc_schwarz(const orc_basis* b, int spherical, double* out) {
    init_tables(b);
    const int ng = b->ngroups;
#pragma omp parallel
    {
        double* cart = (double*)malloc(sizeof(double) * 50625);
        double* blk = (double*)malloc(sizeof(double) * 50625);
        double* tmp = (double*)malloc(sizeof(double) * 2 * 50625);
        int i0[15], i1[15], i2[15], i3[15];
        int* idx[4] = {i0, i1, i2, i3};
#pragma omp for schedule(dynamic)
        for (int gi = 0; gi < ng; ++gi)
            for (int gj = 0; gj <= gi; ++gj) {
                int dims[4];
                out_block(b, spherical, gi, gj, gi, gj, 0.0, cart, blk, tmp, dims, idx);
                double mx = 0.0;
                for (int a = 0; a < dims[0]; ++a)
                    for (int c = 0; c < dims[1]; ++c) {
                        const size_t o = (((size_t)a * dims[1] + c) * dims[2] + a) * dims[3] + c;
                        if (fabs(blk[o]) > mx) mx = fabs(blk[o]);
                    }
                out[gi * (gi + 1) / 2 + gj] = sqrt(mx);
            }
        free(cart);
        free(blk);
        free(tmp);
    }
    return 0;
}

/* Heuristic shell-pair decay indicator (rys_engine.ipp:2205-2241):
 * K_g = sum_i |c_i| (pi/a_i)^{3/4} over a group's radial coefficients,
 * D_ab = K_a K_b exp(-mu_min |AB|^2), mu_min from the two smallest exponents. */
static double decay_scale(const orc_group* g, double* amin) {
    double k = 0.0, a0 = g->exps[0];
    for (int i = 0; i < g->nprim; ++i) {
        if (g->exps[i] < a0) a0 = g->exps[i];
        k += fabs(g->c_radial[i]) * pow(PI_D / g->exps[i], 0.75);
    }
    *amin = a0;
    return k;
}

int orc_jk(const orc_basis* b, int spherical, double schwarz, const double* Dm,
           double* J, double* K, int nthreads) {
    return orc_jk_screened(b, spherical, schwarz, 0.0, 0.0, Dm, J, K, nthreads);
}

/* Shell quartets: with schwarz > 0, kept when Q_ab Q_cd >= schwarz and (decay
 * > 0) D_ab D_cd >= decay; with schwarz == 0 and decay > 0, kept when
 * D_ab D_cd >= decay (native_quartets.ipp:690-705, 49-53). */
int orc_jk_screened(const orc_basis* b, int spherical, double schwarz, double decay, double fastschwarz,
                    const double* Dm, double* J, double* K, int nthreads) {
    init_tables(b);
    const double fs2 = fastschwarz * fastschwarz;
    const int n = spherical ? b->nsph : b->ncart;
    const int ng = b->ngroups;
    const int npair = ng * (ng + 1) / 2;
    double* Q = (double*)malloc(sizeof(double) * (size_t)npair);
    if (schwarz > 0.0) orc_schwarz(b, spherical, Q);
    double* Dc = (double*)malloc(sizeof(double) * (size_t)npair);
    if (decay > 0.0)
        for (int gi = 0, p = 0; gi < ng; ++gi)
            for (int gj = 0; gj <= gi; ++gj, ++p) {
                double ai, aj;
                const double ki = decay_scale(&b->groups[gi], &ai), kj = decay_scale(&b->groups[gj], &aj);
                double r2 = 0.0;
                for (int x = 0; x < 3; ++x) {
                    const double d = b->groups[gi].center[x] - b->groups[gj].center[x];
                    r2 += d * d;
                }
                Dc[p] = ki * kj * exp(-(ai * aj / (ai + aj)) * r2);
            }
    int* pi = (int*)malloc(sizeof(int) * (size_t)npair);
    int* pj = (int*)malloc(sizeof(int) * (size_t)npair);
    for (int gi = 0, p = 0; gi < ng; ++gi)
        for (int gj = 0; gj <= gi; ++gj, ++p) {
            pi[p] = gi;
            pj[p] = gj;
        }
    const size_t n2 = (size_t)n * n;
    if (J) memset(J, 0, sizeof(double) * n2);
    if (K) memset(K, 0, sizeof(double) * n2);
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    const int nt = 1;
    (void)nthreads;
#endif
    double* Jt = (double*)calloc((size_t)nt * n2, sizeof(double));
    double* Kt = (double*)calloc((size_t)nt * n2, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        double* Jl = Jt + (size_t)tid * n2;
        double* Kl = Kt + (size_t)tid * n2;
        double* cart = (double*)malloc(sizeof(double) * 50625);
        double* blk = (double*)malloc(sizeof(double) * 50625);
        double* tmp = (double*)malloc(sizeof(double) * 2 * 50625);
        int i0[15], i1[15], i2[15], i3[15];
        int* idx[4] = {i0, i1, i2, i3};
#pragma omp for schedule(dynamic, 1)
        for (int P = 0; P < npair; ++P)
            for (int R = 0; R <= P; ++R) {
                if (schwarz > 0.0 && Q[P] * Q[R] < schwarz) continue;
                if (decay > 0.0 && Dc[P] * Dc[R] < decay) continue;
                const int g[4] = {pi[P], pj[P], pi[R], pj[R]};
                int dims[4];
                out_block(b, spherical, g[0], g[1], g[2], g[3], fs2, cart, blk, tmp, dims, idx);
                /* the 8 orderings as positions of (i,j,k,l) */
                static const int perm[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                               {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
                int seen[8][4];
                int nseen = 0;
                for (int s = 0; s < 8; ++s) {
                    int q[4];
                    for (int k = 0; k < 4; ++k) q[k] = g[perm[s][k]];
                    int dup = 0;
                    for (int t = 0; t < nseen; ++t)
                        if (seen[t][0] == q[0] && seen[t][1] == q[1] && seen[t][2] == q[2] && seen[t][3] == q[3])
                            dup = 1;
                    if (dup) continue;
                    memcpy(seen[nseen++], q, sizeof q);
                    const int* pm = perm[s];
                    /* iterate the block with roles permuted */
                    int e[4];
                    for (e[0] = 0; e[0] < dims[0]; ++e[0])
                        for (e[1] = 0; e[1] < dims[1]; ++e[1])
                            for (e[2] = 0; e[2] < dims[2]; ++e[2])
                                for (e[3] = 0; e[3] < dims[3]; ++e[3]) {
                                    const double v =
                                        blk[(((size_t)e[0] * dims[1] + e[1]) * dims[2] + e[2]) * dims[3] + e[3]];
                                    const int a = idx[pm[0]][e[pm[0]]];
                                    const int bb = idx[pm[1]][e[pm[1]]];
                                    const int c = idx[pm[2]][e[pm[2]]];
                                    const int d = idx[pm[3]][e[pm[3]]];
                                    if (J) Jl[(size_t)a * n + bb] += v * Dm[(size_t)c * n + d];
                                    if (K) Kl[(size_t)a * n + c] += v * Dm[(size_t)bb * n + d];
                                }
                }
            }
        free(cart);
        free(blk);
        free(tmp);
    }
    for (int t = 0; t < nt; ++t)
        for (size_t x = 0; x < n2; ++x) {
            if (J) J[x] += Jt[(size_t)t * n2 + x];
            if (K) K[x] += Kt[(size_t)t * n2 + x];
        }
    free(Jt);
    free(Kt);
    free(Q);
    free(Dc);
    free(pi);
    free(pj);
    return 0;
}
