/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement in plain C of the reference's Gaussian-integral path
 * (Boys function, Rys roots/weights, Rys-quadrature ERIs with HRR, direct
 * J/K digestion, one-electron S/T/V).  It is the checker for the HIP
 * library: only tests, the smoke entry and the CPU-baseline leg of the
 * benchmark load it, and the product path never calls into it.
 *
 * Every routine is written from the mathematics, following the reference's
 * conventions, which are cited per function in the .c files (paths relative
 * to the reference root).  Pinning of each routine (golden vectors from the
 * compiled reference, closed forms, identities) lives in tests/.
 *
 * Basis input: an array of Cartesian COMPONENTS exactly as the reference's
 * std::vector<Shell> (core/basis/types.hpp:31-36): component i has centre
 * centers[3i..3i+2] (bohr), powers ang[3i..3i+2], nprim[i] primitives with
 * consecutive exponents/coefficients (coefficients already normalized,
 * core/shell/shell_contract.cpp:9-48).
 */
#ifndef ORACLE_H
#define ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boys F_m(x), long-double series / erf + upward recursion. */
double orc_boys(int m, double x);
void orc_boys_all(int mmax, double x, double* out);

/* Rys rule of order N for argument T: roots t^2 in [0,1), weights; exact for
 * int_0^1 f(t^2) exp(-T t^2) dt with deg f <= 2N-1.  Returns 0 on success. */
int orc_rys(int N, double T, double* roots, double* weights);

/* normalize_contracted_primitives restated: raw -> stored coefficients. */
void orc_normalize(int nprim, const double* exps, const double* coefs,
                   const int* ang, double* out);

/* Real-solid-harmonic projection for shell L (unit-diagonal Cartesian input):
 * fills c2s[(2L+1) x ncart(L)] row-major, spherical rows in PySCF order. */
void orc_cart2sph(int L, double* c2s);

typedef struct orc_basis orc_basis;

orc_basis* orc_basis_create(int n, const double* centers, const int* ang,
                            const int* nprim, const double* exps,
                            const double* coefs);
void orc_basis_destroy(orc_basis* b);
const char* orc_last_error(void);
int orc_ngroups(const orc_basis* b);
int orc_ncart(const orc_basis* b);
int orc_nsph(const orc_basis* b);
/* group g: L, component basis indices (ncart(L) entries), spherical base */
int orc_group_info(const orc_basis* b, int g, int* L, int* comps, int* sph_base);

/* Unit-diagonal Cartesian block (ga gb|gc gd), [a][b][c][d] in component order. */
int orc_eri_cart_block(const orc_basis* b, int ga, int gb, int gc, int gd, double* out);

/* Direct J and K (either may be NULL).  spherical = 0: Cartesian AO basis
 * (component order, n = ncart); 1: spherical basis (M = nsph).  Shell
 * quartets with Q_ab * Q_cd < schwarz are skipped (0 disables).  D is
 * symmetric n x n row-major.  nthreads <= 0 uses the OpenMP default. */
int orc_jk(const orc_basis* b, int spherical, double schwarz, const double* D,
           double* J, double* K, int nthreads);

/* orc_jk with the reference's heuristic screens: decay > 0 adds the shell-pair
 * decay indicator test (alone when schwarz == 0), fastschwarz > 0 the
 * primitive FASTSCHWARZ test on every primitive quartet. */
int orc_jk_screened(const orc_basis* b, int spherical, double schwarz, double decay, double fastschwarz,
                    const double* D, double* J, double* K, int nthreads);

/* Schwarz pair bound per group pair (gi >= gj, index gi*(gi+1)/2+gj). */
int orc_schwarz(const orc_basis* b, int spherical, double* out);

/* One-electron S, T, V; nuclear charges Z and centres R (3*n_nuc). */
int orc_onee(const orc_basis* b, int spherical, int n_nuc, const double* Z,
             const double* R, double* S, double* T, double* V);

#ifdef __cplusplus
}
#endif
#endif
