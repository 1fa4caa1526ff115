// Table generation for the device Boys / Rys evaluators (see tables.hpp).
#include "tables.hpp"

#include <cmath>
#include <cstring>
#include <mutex>

namespace tlno_amd {
namespace {

typedef long double ld;
constexpr ld kPiL = 3.141592653589793238462643383279502884L;

ld boys_ld(int m, ld x) {
    if (x <= 0.0L) return 1.0L / (2.0L * m + 1.0L);
    if (x < 100.0L || ld(m) > x - 10.0L) {
        ld term = 1.0L / (2.0L * m + 1.0L), sum = term;
        for (int k = 1; k < 4000; ++k) {
            term *= 2.0L * x / (2.0L * m + 2.0L * k + 1.0L);
            sum += term;
            if (term < 1e-22L * sum) break;
        }
        return expl(-x) * sum;
    }
    ld f = 0.5L * sqrtl(kPiL / x) * erfl(sqrtl(x));
    const ld ex = expl(-x);
    for (int k = 0; k < m; ++k) f = ((2.0L * k + 1.0L) * f - ex) / (2.0L * x);
    return f;
}

constexpr int kGL = 320;
struct GaussLegendre01 {
    ld x[kGL], w[kGL];
    GaussLegendre01() {
        for (int i = 0; i < (kGL + 1) / 2; ++i) {
            ld z = cosl(kPiL * (i + 0.75L) / (kGL + 0.5L)), dp = 0.0L;
            for (int it = 0; it < 100; ++it) {
                ld p1 = 1.0L, p2 = 0.0L;
                for (int j = 1; j <= kGL; ++j) {
                    const ld p3 = p2;
                    p2 = p1;
                    p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
                }
                dp = kGL * (z * p1 - p2) / (z * z - 1.0L);
                const ld dz = p1 / dp;
                z -= dz;
                if (fabsl(dz) < 1e-21L) break;
            }
            const ld wt = 1.0L / ((1.0L - z * z) * dp * dp);  // half of 2/(...)
            x[i] = 0.5L * (1.0L - z);
            x[kGL - 1 - i] = 0.5L * (1.0L + z);
            w[i] = wt;
            w[kGL - 1 - i] = wt;
        }
    }
};

// Implicit QL on a symmetric tridiagonal matrix; d diagonal, e[i] couples
// i and i+1 (e[n-1] unused).  Returns eigenvalues in d and the first row of
// the eigenvector matrix in z.
void tridiag_ql(int n, ld* d, ld* e, ld* z) {
    ld Z[20][20];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z[i][j] = (i == j) ? 1.0L : 0.0L;
    e[n - 1] = 0.0L;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const ld dd = fabsl(d[m]) + fabsl(d[m + 1]);
                if (fabsl(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (++iter > 300) break;
            ld g = (d[l + 1] - d[l]) / (2.0L * e[l]);
            ld r = hypotl(g, 1.0L);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            ld s = 1.0L, c = 1.0L, p = 0.0L;
            int i;
            bool underflow = false;
            for (i = m - 1; i >= l; --i) {
                ld f = s * e[i];
                const ld b = c * e[i];
                r = hypotl(f, g);
                e[i + 1] = r;
                if (r == 0.0L) {
                    d[i + 1] -= p;
                    e[m] = 0.0L;
                    underflow = true;
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
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0L;
        } while (true);
    }
    for (int j = 0; j < n; ++j) z[j] = Z[0][j];
}

void sort_pairs(int n, ld* a, ld* b) {
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (a[j] < a[i]) {
                ld t = a[i]; a[i] = a[j]; a[j] = t;
                t = b[i]; b[i] = b[j]; b[j] = t;
            }
}

void hermite_half(int N, ld* r2, ld* w) {
    const int M = 2 * N;
    ld d[20], e[20], z[20];
    for (int i = 0; i < M; ++i) {
        d[i] = 0.0L;
        e[i] = (i + 1 < M) ? sqrtl(0.5L * (i + 1)) : 0.0L;
    }
    tridiag_ql(M, d, e, z);
    ld zz[20];
    for (int i = 0; i < M; ++i) zz[i] = z[i] * z[i];
    sort_pairs(M, d, zz);
    for (int i = 0; i < N; ++i) {
        r2[i] = d[N + i] * d[N + i];
        w[i] = sqrtl(kPiL) * zz[N + i];
    }
}

void rys_ld(int N, ld T, ld* r, ld* w) {
    static const GaussLegendre01 gl;
    ld xs[kGL], ws[kGL], p0[kGL], p1[kGL];
    ld alpha[20], beta[20];
    ld s0 = 0.0L, s1 = 0.0L;
    for (int j = 0; j < kGL; ++j) {
        xs[j] = gl.x[j] * gl.x[j];
        ws[j] = gl.w[j] * expl(-T * xs[j]);
        s0 += ws[j];
        s1 += ws[j] * xs[j];
        p0[j] = 0.0L;
        p1[j] = 1.0L;
    }
    alpha[0] = s1 / s0;
    beta[0] = s0;
    ld prev = s0;
    for (int k = 0; k + 1 < N; ++k) {
        ld n0 = 0.0L, n1 = 0.0L;
        for (int j = 0; j < kGL; ++j) {
            const ld p2 = (xs[j] - alpha[k]) * p1[j] - (k > 0 ? beta[k] : 0.0L) * p0[j];
            const ld t = ws[j] * p2 * p2;
            n0 += t;
            n1 += t * xs[j];
            p0[j] = p1[j];
            p1[j] = p2;
        }
        alpha[k + 1] = n1 / n0;
        beta[k + 1] = n0 / prev;
        prev = n0;
    }
    ld d[20], e[20], z[20];
    for (int i = 0; i < N; ++i) {
        d[i] = alpha[i];
        e[i] = (i + 1 < N) ? sqrtl(beta[i + 1]) : 0.0L;
    }
    tridiag_ql(N, d, e, z);
    for (int i = 0; i < N; ++i) {
        r[i] = d[i];
        w[i] = beta[0] * z[i] * z[i];
    }
    sort_pairs(N, r, w);
}

HostTables* build_tables() {
    auto* t = new HostTables();
    // Boys
    t->boys.assign(static_cast<std::size_t>(BOYS_NGRID) * (BOYS_MMAX + 1), 0.0);
    for (int k = 0; k < BOYS_NGRID; ++k)
        for (int m = 0; m <= BOYS_MMAX; ++m)
            t->boys[static_cast<std::size_t>(k) * (BOYS_MMAX + 1) + m] = double(boys_ld(m, ld(k) / BOYS_STEPS));
    t->boys_lds.assign(static_cast<std::size_t>(BOYS_LDS_NGRID) * BOYS_LDS_M, 0.0);
    for (int k = 0; k < BOYS_LDS_NGRID; ++k)
        for (int m = 0; m < BOYS_LDS_M; ++m)
            t->boys_lds[static_cast<std::size_t>(k) * BOYS_LDS_M + m] = double(boys_ld(m, ld(k) / BOYS_LDS_STEPS));
    // Rys: Chebyshev interpolation at D+1 nodes per unit interval, converted
    // to monomials in u = T - (k + 1/2), all in long double.
    int off = 0;
    t->rys_off[0] = 0;
    t->rys_off[1] = 0;
    for (int N = 2; N <= RYS_NMAX; ++N) {
        t->rys_off[N] = off;
        off += rys_ttab(N) * 2 * N * RYS_NCOEF;
    }
    t->rys_off[RYS_NMAX + 1] = off;
    t->rys.assign(static_cast<std::size_t>(off), 0.0);
    constexpr int n = RYS_NCOEF;
    ld xnode[n];
    for (int j = 0; j < n; ++j) xnode[j] = cosl(kPiL * (j + 0.5L) / n);
    // Chebyshev T_i(x) as monomials in x: cheb_mono[i][p]
    ld cheb_mono[n][n] = {};
    cheb_mono[0][0] = 1.0L;
    cheb_mono[1][1] = 1.0L;
    for (int i = 2; i < n; ++i)
        for (int p = 0; p < n; ++p)
            cheb_mono[i][p] = (p > 0 ? 2.0L * cheb_mono[i - 1][p - 1] : 0.0L) - cheb_mono[i - 2][p];
    for (int N = 2; N <= RYS_NMAX; ++N) {
        const int nint = rys_ttab(N);
        std::vector<ld> vals(static_cast<std::size_t>(n) * 2 * N);
#pragma omp parallel for schedule(dynamic) firstprivate(vals)
        for (int k = 0; k < nint; ++k) {
            for (int j = 0; j < n; ++j) {
                ld r[20], w[20];
                const ld T = k + 0.5L + 0.5L * xnode[j];  // u in [-1/2, 1/2] <-> x in [-1,1]
                rys_ld(N, T, r, w);
                for (int i = 0; i < N; ++i) {
                    vals[static_cast<std::size_t>(i) * n + j] = r[i];
                    vals[static_cast<std::size_t>(N + i) * n + j] = w[i];
                }
            }
            for (int v = 0; v < 2 * N; ++v) {
                ld c[n];
                for (int i = 0; i < n; ++i) {
                    ld s = 0.0L;
                    for (int j = 0; j < n; ++j)
                        s += vals[static_cast<std::size_t>(v) * n + j] * cosl(kPiL * i * (j + 0.5L) / n);
                    c[i] = (i == 0 ? 1.0L : 2.0L) * s / n;
                }
                // sum_i c_i T_i(x), x = 2u
                for (int p = 0; p < n; ++p) {
                    ld mono = 0.0L;
                    for (int i = 0; i < n; ++i) mono += c[i] * cheb_mono[i][p];
                    mono *= powl(2.0L, p);
                    t->rys[static_cast<std::size_t>(t->rys_off[N]) +
                           (static_cast<std::size_t>(k) * 2 * N + v) * n + p] = double(mono);
                }
            }
        }
    }
    for (int N = 1; N <= RYS_NMAX; ++N) {
        ld r2[20], w[20];
        hermite_half(N, r2, w);
        for (int i = 0; i < N; ++i) {
            t->herm_r2[N][i] = double(r2[i]);
            t->herm_w[N][i] = double(w[i]);
        }
    }
    return t;
}

}  // namespace

const HostTables& host_tables() {
    static std::once_flag once;
    static HostTables* tables = nullptr;
    std::call_once(once, [] { tables = build_tables(); });
    return *tables;
}

void rys_accurate(int N, double T, double* roots, double* weights) {
    ld r[20], w[20];
    rys_ld(N, T, r, w);
    for (int i = 0; i < N; ++i) {
        roots[i] = double(r[i]);
        weights[i] = double(w[i]);
    }
}

double boys_accurate(int m, double T) { return double(boys_ld(m, T)); }

}  // namespace tlno_amd
