/*
 * ORACLE (test infrastructure only): normalization, shell grouping and the
 * Cartesian -> real-spherical projection.
 *
 * Conventions followed (reference paths):
 *  - Primitive normalization and per-component unit self-overlap:
 *    core/shell/shell_math.cpp:50-77, core/shell/shell_contract.cpp:9-48.
 *  - Cartesian order inside a shell is descending lex (x then y):
 *    core/basis/cartesian_shells.hpp:16-28.
 *  - Grouping of components into shells by (L, centre, radial contraction
 *    equal within 32 eps), c_radial = c / N_A(first component),
 *    N_A = 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!), exactly one complete
 *    Cartesian shell per group: integrals/two_electron/rys/rys_engine.ipp:
 *    394-398, 1925-2062.
 *  - Spherical AO numbering: groups in first-appearance order, 2L+1 each
 *    (rys_engine.ipp:2365-2379); PySCF real solid harmonics with the l = 1
 *    order x, y, z (core/shell/cart2sph_tables.hpp:1-5).  The projection is
 *    generated here from the solid-harmonic recursion (Helgaker, Jorgensen,
 *    Olsen, "Molecular Electronic-Structure Theory", eqs. 6.4.70-6.4.73) and
 *    normalized in the unit-diagonal Cartesian metric; it is pinned against
 *    the reference's spherical S and J/K outputs in tests.
 */
#include "oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793238462643383279502884

static __thread char g_err[256];
const char* orc_last_error(void) { return g_err; }
void orc_set_error(const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); }

double orc_dfact(int n) { /* (2n-1)!! */
    double v = 1.0;
    for (int k = 2 * n - 1; k > 0; k -= 2) v *= k;
    return v;
}

int orc_ncart_L(int L) { return (L + 1) * (L + 2) / 2; }

int orc_cart_local(int L, int lx, int ly) {
    const int dx = L - lx;
    return dx * (dx + 1) / 2 + (dx - ly);
}

void orc_cart_tuple(int L, int idx, int* t) {
    int k = 0;
    for (int x = L; x >= 0; --x)
        for (int y = L - x; y >= 0; --y, ++k)
            if (k == idx) {
                t[0] = x; t[1] = y; t[2] = L - x - y;
                return;
            }
}

/* same-centre overlap of one unnormalized Cartesian primitive pair */
static double same_center_overlap(const int* ang, double a, double b) {
    const double p = a + b;
    double f = pow(ORC_PI / p, 1.5);
    for (int k = 0; k < 3; ++k) f *= orc_dfact(ang[k]) / pow(2.0 * p, ang[k]);
    return f;
}

void orc_normalize(int nprim, const double* exps, const double* coefs,
                   const int* ang, double* out) {
    const int L = ang[0] + ang[1] + ang[2];
    const double df = orc_dfact(ang[0]) * orc_dfact(ang[1]) * orc_dfact(ang[2]);
    for (int i = 0; i < nprim; ++i) {
        const double num = pow(2.0, 2.0 * L + 1.5) * pow(exps[i], L + 1.5);
        out[i] = coefs[i] * sqrt(num / (pow(ORC_PI, 1.5) * df));
    }
    double n2 = 0.0;
    for (int i = 0; i < nprim; ++i)
        for (int j = 0; j < nprim; ++j) n2 += out[i] * out[j] * same_center_overlap(ang, exps[i], exps[j]);
    const double s = 1.0 / sqrt(n2);
    for (int i = 0; i < nprim; ++i) out[i] *= s;
}

/* ---- real solid harmonics as monomial polynomials ---- */
/* polynomial of total degree L stored over monomials (ncart(L)) in local order */
typedef struct { double c[64]; } poly_t;

static void poly_zero(poly_t* p) { memset(p, 0, sizeof *p); }

/* multiply degree-L polynomial by x_k -> degree L+1 */
static void poly_mul_axis(const poly_t* a, int L, int axis, double s, poly_t* out) {
    for (int i = 0; i < orc_ncart_L(L); ++i) {
        if (a->c[i] == 0.0) continue;
        int t[3];
        orc_cart_tuple(L, i, t);
        t[axis] += 1;
        out->c[orc_cart_local(L + 1, t[0], t[1])] += s * a->c[i];
    }
}

/* multiply degree-L polynomial by r^2 -> degree L+2 */
static void poly_mul_r2(const poly_t* a, int L, double s, poly_t* out) {
    for (int i = 0; i < orc_ncart_L(L); ++i) {
        if (a->c[i] == 0.0) continue;
        int t[3];
        orc_cart_tuple(L, i, t);
        for (int k = 0; k < 3; ++k) {
            int u[3] = {t[0], t[1], t[2]};
            u[k] += 2;
            out->c[orc_cart_local(L + 2, u[0], u[1])] += s * a->c[i];
        }
    }
}

void orc_cart2sph(int Lmax_req, double* c2s) {
    /* S[l][m + l] for l <= Lmax_req, recursion (6.4.70-6.4.73). */
    enum { LM = 6 };
    static poly_t S[LM + 1][2 * LM + 1];
    const int Lmax = Lmax_req;
    for (int l = 0; l <= Lmax; ++l)
        for (int m = 0; m <= 2 * l; ++m) poly_zero(&S[l][m]);
    S[0][0].c[0] = 1.0;
    for (int l = 0; l < Lmax; ++l) {
        const double f = sqrt((l == 0 ? 2.0 : 1.0) * (2.0 * l + 1.0) / (2.0 * l + 2.0));
        /* S_{l+1,l+1} = f (x S_{l,l} - (1-d_l0) y S_{l,-l}) */
        poly_mul_axis(&S[l][2 * l], l, 0, f, &S[l + 1][2 * l + 2]);
        if (l > 0) poly_mul_axis(&S[l][0], l, 1, -f, &S[l + 1][2 * l + 2]);
        /* S_{l+1,-l-1} = f (y S_{l,l} + (1-d_l0) x S_{l,-l}) */
        poly_mul_axis(&S[l][2 * l], l, 1, f, &S[l + 1][0]);
        if (l > 0) poly_mul_axis(&S[l][0], l, 0, f, &S[l + 1][0]);
        /* S_{l+1,m} = ((2l+1) z S_{lm} - sqrt((l+m)(l-m)) r^2 S_{l-1,m}) / sqrt((l+m+1)(l-m+1)) */
        for (int m = -l; m <= l; ++m) {
            const double den = sqrt((double)(l + m + 1) * (l - m + 1));
            poly_mul_axis(&S[l][m + l], l, 2, (2.0 * l + 1.0) / den, &S[l + 1][m + l + 1]);
            if (l > 0 && abs(m) <= l - 1)
                poly_mul_r2(&S[l - 1][m + l - 1], l - 1, -sqrt((double)(l + m) * (l - m)) / den,
                            &S[l + 1][m + l + 1]);
        }
    }
    const int L = Lmax;
    const int nc = orc_ncart_L(L);
    const int ns = 2 * L + 1;
    for (int s = 0; s < ns; ++s) {
        int m;
        if (L == 1) m = (s == 0) ? 1 : (s == 1 ? -1 : 0); /* x, y, z */
        else m = s - L;
        const poly_t* P = &S[L][m + L];
        /* coefficient on the unit-diagonal function: mono coef * sqrt(df) */
        double row[64];
        for (int c = 0; c < nc; ++c) {
            int t[3];
            orc_cart_tuple(L, c, t);
            row[c] = P->c[c] * sqrt(orc_dfact(t[0]) * orc_dfact(t[1]) * orc_dfact(t[2]));
        }
        /* norm in the unit-diagonal metric: <phi_c|phi_d> = df(c+d)/sqrt(df(c)df(d)) per axis */
        double n2 = 0.0;
        for (int c = 0; c < nc; ++c) {
            int tc[3];
            orc_cart_tuple(L, c, tc);
            for (int d = 0; d < nc; ++d) {
                int td[3];
                orc_cart_tuple(L, d, td);
                double g = 1.0;
                for (int k = 0; k < 3; ++k) {
                    const int sum = tc[k] + td[k];
                    if (sum % 2) { g = 0.0; break; }
                    g *= orc_dfact(sum / 2) /
                         sqrt(orc_dfact(tc[k]) * orc_dfact(td[k]));
                }
                n2 += row[c] * row[d] * g;
            }
        }
        const double scale = 1.0 / sqrt(n2);
        for (int c = 0; c < nc; ++c) c2s[s * nc + c] = row[c] * scale;
    }
}

/* ---------------- basis / groups ---------------- */

static int same_radial(const orc_group* g, const double* exps, const double* coefs,
                       int np, const int* ang) {
    if (g->nprim != np) return 0;
    for (int i = 0; i < np; ++i)
        if (g->exps[i] != exps[i]) return 0;
    const double sn = 1.0 / sqrt(orc_dfact(ang[0]) * orc_dfact(ang[1]) * orc_dfact(ang[2]));
    for (int i = 0; i < np; ++i) {
        const double a = g->c_radial[i];
        const double b = coefs[i] / sn;
        if (a == b) continue;
        const double sc = fmax(fabs(a), fabs(b));
        if (sc == 0.0) continue;
        if (fabs(a - b) > 32.0 * 2.220446049250313e-16 * sc) return 0;
    }
    return 1;
}

orc_basis* orc_basis_create(int n, const double* centers, const int* ang,
                            const int* nprim, const double* exps,
                            const double* coefs) {
    orc_rys_init();
    orc_basis* b = (orc_basis*)calloc(1, sizeof *b);
    b->ncart = n;
    b->groups = (orc_group*)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_group));
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        const int* a = ang + 3 * i;
        const int L = a[0] + a[1] + a[2];
        const double* ex = exps + off;
        const double* cf = coefs + off;
        const int np = nprim[i];
        off += (size_t)np;
        if (a[0] < 0 || a[1] < 0 || a[2] < 0 || L > 4 || np < 1 || np > ORC_MAXPRIM) {
            orc_set_error("oracle: invalid component");
            orc_basis_destroy(b);
            return NULL;
        }
        int gi = -1;
        for (int g = 0; g < b->ngroups; ++g) {
            orc_group* G = &b->groups[g];
            if (G->L == L && G->center[0] == centers[3 * i] && G->center[1] == centers[3 * i + 1] &&
                G->center[2] == centers[3 * i + 2] && same_radial(G, ex, cf, np, a)) {
                gi = g;
                break;
            }
        }
        if (gi < 0) {
            gi = b->ngroups++;
            orc_group* G = &b->groups[gi];
            G->L = L;
            for (int k = 0; k < 3; ++k) G->center[k] = centers[3 * i + k];
            G->nprim = np;
            const double na = 1.0 / sqrt(orc_dfact(a[0]) * orc_dfact(a[1]) * orc_dfact(a[2]));
            for (int p = 0; p < np; ++p) {
                G->exps[p] = ex[p];
                G->c_radial[p] = cf[p] / na;
            }
        }
        orc_group* G = &b->groups[gi];
        if (G->ncomp >= orc_ncart_L(L)) {
            orc_set_error("oracle: duplicate component in a contracted shell");
            orc_basis_destroy(b);
            return NULL;
        }
        G->comp[G->ncomp] = i;
        for (int k = 0; k < 3; ++k) G->ang[G->ncomp][k] = a[k];
        G->n_ang[G->ncomp] = 1.0 / sqrt(orc_dfact(a[0]) * orc_dfact(a[1]) * orc_dfact(a[2]));
        G->ncomp++;
    }
    int sph = 0;
    for (int g = 0; g < b->ngroups; ++g) {
        orc_group* G = &b->groups[g];
        if (G->ncomp != orc_ncart_L(G->L)) {
            orc_set_error("oracle: each radial contraction must hold one complete Cartesian shell");
            orc_basis_destroy(b);
            return NULL;
        }
        G->sph_base = sph;
        sph += 2 * G->L + 1;
        orc_cart2sph(G->L, G->c2s);
        /* c2s rows are indexed by local Cartesian order; component order may
         * differ, so build the component-order projector too */
        const int nc = G->ncomp;
        for (int s = 0; s < 2 * G->L + 1; ++s)
            for (int c = 0; c < nc; ++c)
                G->c2s_comp[s * nc + c] =
                    G->c2s[s * nc + orc_cart_local(G->L, G->ang[c][0], G->ang[c][1])];
    }
    b->nsph = sph;
    return b;
}

void orc_basis_destroy(orc_basis* b) {
    if (!b) return;
    free(b->groups);
    free(b);
}

int orc_ngroups(const orc_basis* b) { return b->ngroups; }
int orc_ncart(const orc_basis* b) { return b->ncart; }
int orc_nsph(const orc_basis* b) { return b->nsph; }

int orc_group_info(const orc_basis* b, int g, int* L, int* comps, int* sph_base) {
    if (g < 0 || g >= b->ngroups) return 1;
    const orc_group* G = &b->groups[g];
    if (L) *L = G->L;
    if (comps)
        for (int c = 0; c < G->ncomp; ++c) comps[c] = G->comp[c];
    if (sph_base) *sph_base = G->sph_base;
    return 0;
}
