/*
 * ORACLE (test infrastructure only): Boys function and Rys roots/weights.
 *
 * Boys F_m(x) = int_0^1 t^{2m} exp(-x t^2) dt.  The reference tabulates it
 * (core/math/general_math.cpp:18-87 for the 1e engine, 96-point Gauss-
 * Legendre table, Newton interpolation, erf + upward recursion for x >= 12;
 * integrals/two_electron/rys/rys_engine.ipp:492-598 for the 2e copy).  The
 * oracle evaluates the same function directly and more accurately:
 *   x < 100 : F_m(x) = e^{-x} sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1)),
 *             all terms positive, summed in long double;
 *   x >= 100: F_0 = sqrt(pi/x) erf(sqrt x) / 2 and the upward recursion
 *             F_{m+1} = ((2m+1) F_m - e^{-x}) / (2x), stable for m < x.
 *
 * Rys rule (rys_engine.ipp:3366-3739): roots t_i^2 and weights w_i with
 * sum_i w_i t_i^{2k} = F_k(T) for k < 2N.  The reference seeds its tables by a
 * discretized Stieltjes procedure on a Fejer grid followed by a Jacobi
 * eigen-solve of the recurrence matrix (Golub-Welsch), and switches to the
 * half-range Hermite asymptote r_i^2/T, w_i/sqrt(T) above
 * T_asym(N) = {29,37,43,49,55,60,65,71,76} (rys_engine.ipp:3543-3547).  The
 * oracle runs the same Stieltjes + Golub-Welsch construction per call in long
 * double on a 320-point Gauss-Legendre grid in t (exact for the polynomial
 * part, resolving exp(-T t^2) for the T range used), and uses the Hermite
 * asymptote only for T >= 400 where it is exact to binary64.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793238462643383279502884L

static long double boys_series(int m, long double x) {
    long double term = 1.0L / (2.0L * m + 1.0L);
    long double sum = term;
    for (int k = 1; k < 4000; ++k) {
        term *= 2.0L * x / (2.0L * m + 2.0L * k + 1.0L);
        sum += term;
        if (term < 1e-22L * sum) break;
    }
    return expl(-x) * sum;
}

static long double boys_ld(int m, long double x) {
    if (x <= 0.0L) return 1.0L / (2.0L * m + 1.0L);
    if (x < 100.0L || (long double)m > x - 10.0L) return boys_series(m, x);
    long double f = 0.5L * sqrtl(ORC_PI / x) * erfl(sqrtl(x));
    const long double ex = expl(-x);
    for (int k = 0; k < m; ++k) f = ((2.0L * k + 1.0L) * f - ex) / (2.0L * x);
    return f;
}

double orc_boys(int m, double x) { return (double)boys_ld(m, (long double)x); }

void orc_boys_all(int mmax, double x, double* out) {
    for (int m = 0; m <= mmax; ++m) out[m] = orc_boys(m, x);
}

/* ---- Gauss-Legendre rule on [0,1] in long double (Newton on P_n). ---- */
#define GL_N 320
static long double gl_x[GL_N], gl_w[GL_N];
static int gl_ready = 0;

static void gl_init(void) {
    if (gl_ready) return;
    const int n = GL_N;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double z = cosl(ORC_PI * (i + 0.75L) / (n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= n; ++j) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            dp = n * (z * p1 - p2) / (z * z - 1.0L);
            long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-21L) break;
        }
        long double w = 2.0L / ((1.0L - z * z) * dp * dp);
        /* map [-1,1] -> [0,1] */
        gl_x[i] = 0.5L * (1.0L - z);
        gl_x[n - 1 - i] = 0.5L * (1.0L + z);
        gl_w[i] = 0.5L * w;
        gl_w[n - 1 - i] = 0.5L * w;
    }
    __atomic_store_n(&gl_ready, 1, __ATOMIC_RELEASE);
}

/* Symmetric tridiagonal eigen-solve (implicit QL with Wilkinson shifts),
 * d[0..n-1] diagonal, e[1..n-1] sub-diagonal; z receives the first row of
 * the eigenvector matrix.  Eigenvalues returned in d. */
static void tql_first_row(int n, long double* d, long double* e, long double* z) {
    long double Z[32][32];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z[i][j] = (i == j) ? 1.0L : 0.0L;
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0L;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; ++m) {
                long double dd = fabsl(d[m]) + fabsl(d[m + 1]);
                if (fabsl(e[m]) + dd == dd) break;
            }
            if (m != l) {
                if (++iter > 200) break;
                long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
                long double r = hypotl(g, 1.0L);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabsl(r) : -fabsl(r)));
                long double s = 1.0L, c = 1.0L, p = 0.0L;
                int i;
                for (i = m - 1; i >= l; --i) {
                    long double f = s * e[i];
                    long double b = c * e[i];
                    r = hypotl(f, g);
                    e[i + 1] = r;
                    if (r == 0.0L) {
                        d[i + 1] -= p;
                        e[m] = 0.0L;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0L * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = Z[k][i + 1];
                        Z[k][i + 1] = s * Z[k][i] + c * f;
                        Z[k][i] = c * Z[k][i] - s * f;
                    }
                }
                if (r == 0.0L && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0L;
            }
        } while (m != l);
    }
    for (int j = 0; j < n; ++j) z[j] = Z[0][j];
}

static void hermite_half_rule(int N, long double T, long double* r, long double* w) {
    /* Gauss-Hermite of order 2N (weight exp(-u^2) on the real line); the N
     * positive nodes give the half-range rule for exp(-T t^2) on [0, inf). */
    const int M = 2 * N;
    long double d[32], e[32], z[32];
    for (int i = 0; i < M; ++i) d[i] = 0.0L;
    e[0] = 0.0L;
    for (int i = 1; i < M; ++i) e[i] = sqrtl(0.5L * i);
    tql_first_row(M, d, e, z);
    /* sort ascending */
    for (int i = 0; i < M; ++i)
        for (int j = i + 1; j < M; ++j)
            if (d[j] < d[i]) {
                long double t = d[i]; d[i] = d[j]; d[j] = t;
                t = z[i]; z[i] = z[j]; z[j] = t;
            }
    const long double mu0 = sqrtl(ORC_PI);
    for (int i = 0; i < N; ++i) {
        const long double u = d[N + i];
        r[i] = u * u / T;
        w[i] = mu0 * z[N + i] * z[N + i] / sqrtl(T);
    }
}

int orc_rys(int N, double Td, double* roots, double* weights) {
    if (N < 1 || N > 16) return 1;
    const long double T = Td;
    long double r[32], w[32];
    if (T >= 400.0L) {
        hermite_half_rule(N, T, r, w);
    } else {
        gl_init();
        /* discretized Stieltjes: nodes x_j = t_j^2, weights gl_w exp(-T t^2) */
        static __thread long double xs[GL_N], ws[GL_N], p0[GL_N], p1[GL_N], p2[GL_N];
        long double alpha[32], beta[32];
        long double s0 = 0.0L, s1 = 0.0L;
        for (int j = 0; j < GL_N; ++j) {
            xs[j] = gl_x[j] * gl_x[j];
            ws[j] = gl_w[j] * expl(-T * xs[j]);
            s0 += ws[j];
            s1 += ws[j] * xs[j];
            p0[j] = 0.0L;
            p1[j] = 1.0L;
        }
        alpha[0] = s1 / s0;
        beta[0] = s0;
        long double norm_prev = s0;
        for (int k = 0; k < N - 1; ++k) {
            long double n0 = 0.0L, n1 = 0.0L;
            for (int j = 0; j < GL_N; ++j) {
                p2[j] = (xs[j] - alpha[k]) * p1[j] - (k > 0 ? beta[k] : 0.0L) * p0[j];
                const long double t = ws[j] * p2[j] * p2[j];
                n0 += t;
                n1 += t * xs[j];
                p0[j] = p1[j];
                p1[j] = p2[j];
            }
            alpha[k + 1] = n1 / n0;
            beta[k + 1] = n0 / norm_prev;
            norm_prev = n0;
        }
        long double d[32], e[32], z[32];
        for (int i = 0; i < N; ++i) {
            d[i] = alpha[i];
            e[i] = i > 0 ? sqrtl(beta[i]) : 0.0L;
        }
        tql_first_row(N, d, e, z);
        for (int i = 0; i < N; ++i) {
            r[i] = d[i];
            w[i] = beta[0] * z[i] * z[i];
        }
        for (int i = 0; i < N; ++i)
            for (int j = i + 1; j < N; ++j)
                if (r[j] < r[i]) {
                    long double t = r[i]; r[i] = r[j]; r[j] = t;
                    t = w[i]; w[i] = w[j]; w[j] = t;
                }
    }
    for (int i = 0; i < N; ++i) {
        roots[i] = (double)r[i];
        weights[i] = (double)w[i];
    }
    return 0;
}

void orc_rys_init(void) { gl_init(); }

/* ---- Tabulated rule for bulk use (J/K over many quartets) ----
 * Roots/weights of orc_rys sampled on T_k = k h (h = 1/256) for
 * T < 80 and interpolated with a 6-point Lagrange stencil (error below
 * 1e-15 relative for N <= 9; checked against orc_rys in tests).  Beyond the
 * table the Hermite asymptote of order N is used, as the reference does
 * above T_asym(N) (rys_engine.ipp:3543-3547, 3753-3837). */
#define RT_H (1.0 / 256.0)
#define RT_TMAX 84.0
#define RT_NPTS (84 * 256 + 8)
static double* rt_root[10];
static double* rt_wt[10];
static long double herm_r2[10][10], herm_w[10][10];

int orc_rys_table_init(int N) {
    if (N < 1 || N > 9) return 1;
    if (rt_root[N]) return 0;
    gl_init();
    double* r = (double*)malloc(sizeof(double) * (size_t)RT_NPTS * N);
    double* w = (double*)malloc(sizeof(double) * (size_t)RT_NPTS * N);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < RT_NPTS; ++k) orc_rys(N, k * RT_H, r + (size_t)k * N, w + (size_t)k * N);
    long double hr[32], hw[32];
    hermite_half_rule(N, 1.0L, hr, hw);
    for (int i = 0; i < N; ++i) {
        herm_r2[N][i] = hr[i];
        herm_w[N][i] = hw[i];
    }
    rt_wt[N] = w;
    __atomic_store_n(&rt_root[N], r, __ATOMIC_RELEASE);
    return 0;
}

void orc_rys_fast(int N, double T, double* roots, double* weights) {
    if (N > 9) {
        orc_rys(N, T, roots, weights);
        return;
    }
    if (T >= RT_TMAX - 4.0) {
        const double iT = 1.0 / T, isT = sqrt(iT);
        for (int i = 0; i < N; ++i) {
            roots[i] = (double)herm_r2[N][i] * iT;
            weights[i] = (double)herm_w[N][i] * isT;
        }
        return;
    }
    const double u = T / RT_H;
    int k = (int)u - 2;
    if (k < 0) k = 0;
    const double t = u - k; /* stencil nodes at 0..5 */
    double L[6];
    for (int j = 0; j < 6; ++j) {
        double v = 1.0;
        for (int m = 0; m < 6; ++m)
            if (m != j) v *= (t - m) / (double)(j - m);
        L[j] = v;
    }
    const double* r = rt_root[N] + (size_t)k * N;
    const double* w = rt_wt[N] + (size_t)k * N;
    for (int i = 0; i < N; ++i) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < 6; ++j) {
            a += L[j] * r[j * N + i];
            b += L[j] * w[j * N + i];
        }
        roots[i] = a;
        weights[i] = b;
    }
}
