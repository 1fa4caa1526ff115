/*
 * tlno_gpu.h — C ABI of the MI355X (gfx950) Gaussian-integral engine.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * nonzero status on failure; tlno_gpu_last_error() then holds a message that
 * uses the reference engine's wording (prefixes "eri_ao_tensor_hgp:" and
 * "jk_direct:").  Host-pointer entry points copy in and out; *_device entry
 * points take device pointers that are already resident in HBM.
 *
 * Basis input is the reference's component layout
 * (core/basis/types.hpp:31-36, std::vector<Shell>): component i has centre
 * centers[3i..3i+2] in bohr, Cartesian powers ang[3i..3i+2], nprim[i]
 * primitives whose exponents and stored (normalized) coefficients are
 * consecutive in exps/coefs.  Components are grouped into shells exactly as
 * the reference does (integrals/two_electron/rys/rys_engine.ipp:1966-2062).
 *
 * AO conventions:
 *   Cartesian mode: the AO index of a function is its component index.
 *   Spherical mode: shells in first-appearance order, 2L+1 real solid
 *   harmonics each in PySCF order (rys_engine.ipp:2365-2379), exactly the
 *   M x M basis of the reference's J/K endpoints.
 */
#ifndef TLNO_GPU_H
#define TLNO_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tlno_gpu_plan tlno_gpu_plan;

#define TLNO_GPU_J 1
#define TLNO_GPU_K 2

/* Message of the last failing call on this thread. */
const char* tlno_gpu_last_error(void);

/* Selects the HIP device for subsequent plans of this thread and builds the
 * device Boys/Rys tables (no reference counterpart: device selection). */
int tlno_gpu_init(int device);

/* Replaces build_direct_plan / build_eri_packed_plan_for_G_rys
 * (integrals/two_electron/fock/jk_direct.hpp:54-56,
 *  integrals/two_electron/twoe_integrals.hpp:104-106).
 * spherical = 1 builds the reference's spherical AO basis, 0 the Cartesian
 * component basis.  Schwarz bounds are computed here, on the GPU. */
int tlno_gpu_plan_create(int n_comp, const double* centers, const int* ang,
                         const int* nprim, const double* exps, const double* coefs,
                         int spherical, tlno_gpu_plan** out);
void tlno_gpu_plan_destroy(tlno_gpu_plan* plan);

/* M (output AO count), shell groups, shell pairs, primitive pairs
 * (EriPackedPlanProfile, twoe_integrals.hpp:68-89). */
int tlno_gpu_plan_info(const tlno_gpu_plan* plan, int* M, int* n_groups,
                       int* n_pairs, long long* n_prim_pairs);

/* Shell groups of the plan in first-appearance order: L[g] and, for each
 * group, the basis (component) index of its Cartesian components in
 * descending-lex order, comps[15*g + k] (unused slots -1).  max_groups bounds
 * both arrays.  Counterpart of the grouping in build_groups_h
 * (rys_engine.ipp:1966-2062). */
int tlno_gpu_plan_groups(const tlno_gpu_plan* plan, int max_groups, int* L, int* comps);

/* Shell pairs of the plan in the plan's order (class ss, ps, pp, ds, dp, dd,
 * fs, fp, fd, ff, gs, gp, gd, gf, gg, then descending Schwarz bound):
 * bra/ket-side group of each pair (higher L first), Schwarz bound Q (sqrt
 * max|(ab|ab)| in the plan's basis), and the first pair index of each class,
 * class_begin[0..15] (class_begin[15] = number of pairs).  Any pointer may be
 * NULL.  Counterpart of the reference's pair-group list and pair_group_max
 * (rys_engine.ipp:2436-2460, 16164-16258). */
int tlno_gpu_plan_pairs(const tlno_gpu_plan* plan, int* group_a, int* group_b, double* Q,
                        int* class_begin);

/* max |(ab|cd)| over the unit-diagonal Cartesian block of each shell quartet
 * (pair_p[i] | pair_q[i]) given as plan pair indices (training labels for the
 * learned screen; no reference counterpart). */
int tlno_gpu_quartet_maxabs(tlno_gpu_plan* plan, int n, const int* pair_p, const int* pair_q,
                            double* out);

/* Replaces build_coulomb_j / build_exchange_k / build_rhf_g
 * (jk_direct.hpp:80-96) and build_jk_direct_for_G_rys (eri_direct_jk.hpp:36-39):
 * J[a,b] = sum_cd P[c,d](ab|cd), K[a,c] = sum_bd P[b,d](ab|cd), G = J - K/2.
 * density, J, K are M x M row-major host arrays; J or K may be NULL.
 * schwarz: shell quartets with Q_ab Q_cd < schwarz are skipped (>= 0).
 * The density must be finite and symmetric (128 eps), as in the reference
 * (fock/detail/direct_build.ipp:29-97). */
int tlno_gpu_build_jk(tlno_gpu_plan* plan, const double* density, double* J,
                      double* K, double schwarz);

/* Same build on device pointers (M x M doubles each, J or K may be NULL),
 * enqueued on the plan's stream and completed before return.  No density
 * validation is done on this path. */
int tlno_gpu_build_jk_device(tlno_gpu_plan* plan, const double* d_density,
                             double* d_J, double* d_K, double schwarz);

/* Allocation-free variant of tlno_gpu_build_jk_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and screening/precision settings; the build's
 * internal-order density, J/K accumulators and counters live in a static
 * device array of the library (one per GPU, 24 MiB + 4 KiB, held until the
 * process ends), so the plan may have at most 1024 Cartesian functions
 * (larger plans are rejected with an error).  The call allocates nothing:
 * the plan's work lists (per Schwarz threshold), routing buffers, team
 * quartet lists, the team working buffer of plans with g shells and the
 * device's side streams must already exist, i.e. one
 * tlno_gpu_build_jk_device (or tlno_gpu_build_jk) call on this plan with the
 * same threshold and settings has run; otherwise it returns an error and
 * does no work.  All calls on one GPU share its array: calls that overlap in
 * time on one GPU (several host threads, several plans) race and must be
 * serialized by the caller. */
int tlno_gpu_build_jk_device_noalloc(tlno_gpu_plan* plan, const double* d_density,
                                     double* d_J, double* d_K, double schwarz);

/* Density-coupled screening, EriPackedBuildOptions::density_threshold /
 * ScreeningOptions::density_threshold (twoe_integrals.hpp:58-64,
 * jk_direct.hpp:18-24): following builds also skip shell quartets with
 * Q_ab Q_cd * 4 max(g(ab), g_max(class of cd)) < density_threshold, where g(x) is
 * the largest |D| in the rows of the shells of pair x.  That factor bounds the
 * reference's CVHF factor max(4 max(rho_ab, rho_cd), cross blocks) from above,
 * so no quartet the reference keeps is dropped.  0 turns it off (default). */
int tlno_gpu_set_density_threshold(tlno_gpu_plan* plan, double density_threshold);

/* The reference's heuristic screens, ScreeningOptions::
 * primitive_fastschwarz_threshold and ::decay_threshold
 * (integrals/two_electron/fock/jk_direct.hpp:18-24,
 * integrals/two_electron/twoe_integrals.hpp:41-66), for following J/K builds:
 *   decay_threshold > 0: shell quartets with D_ab D_cd < decay_threshold are
 *     skipped, D_ab = K_a K_b exp(-mu_min |AB|^2) the shell-pair decay
 *     indicator (rys/rys_engine.ipp:2205-2241); together with the Schwarz
 *     test when schwarz > 0, alone when schwarz = 0
 *     (rys/native_quartets.ipp:690-705); not combinable with the density
 *     threshold (the build then fails with the reference's message);
 *   primitive_fastschwarz_threshold t > 0: primitive quartets with
 *     fs_ab fs_cd < t^2 are skipped, fs = 24.739... w^2 p^{-1/2}
 *     (native_quartets.ipp:595-600).  This engine tests every primitive
 *     quartet; the reference tests a SIMD packet of bra pairs together and
 *     keeps a primitive quartet when any pair of the packet passes, so it
 *     skips a subset of what this engine skips (the two agree where every
 *     packet holds one pair).
 * 0 turns a screen off (the default).  The stream checksum ignores both. */
int tlno_gpu_set_primitive_screens(tlno_gpu_plan* plan, double primitive_fastschwarz_threshold,
                                   double decay_threshold);

/* Canonical ERI stream checksum: the sum of every unique pure-spherical
 * (ab|cd) (a >= b, c >= d, pair(ab) >= pair(cd)) over shell quartets with
 * Q_ab Q_cd >= schwarz, the one-element result of the reference's
 * build_eri_packed_for_G_rys (twoe_integrals.hpp:108-113,
 * rys_engine.ipp:17096-17103).  Spherical plans only. */
int tlno_gpu_eri_checksum(tlno_gpu_plan* plan, double schwarz, double* checksum);

/* Indexed stream fingerprint of the same canonical values, as the
 * reference's TLNO_VALIDATE_RYS_STREAM_FINGERPRINT build accumulates it
 * (rys_engine.ipp:87-105, 270-300; EriStreamFP) and returns it
 * (native_quartets.ipp:759-765; eri_direct_jk.hpp:25-33):
 *   out[7] = {sum x, sum |x|, sum x^2, m1, m2, m3, max |x|},
 *   m_k = sum x w_k(i), i = qidx(pair(a,b), pair(c,d)) the canonical index,
 *   w_k(i) = ((splitmix64(i ^ seed_k) >> 40) - 2^23) / 2^23 with the
 *   reference's seeds; *count = number of canonical values.  Unlike the
 * checksum it changes when a value lands in the wrong slot.  Spherical
 * plans only; allocates scratch for the Cartesian blocks (up to 256 MiB). */
int tlno_gpu_eri_fingerprint(tlno_gpu_plan* plan, double schwarz, double* out, unsigned long long* count);

/* Learned screen (no reference counterpart).  params: TLNO_GPU_MLP_NALL
 * floats of a network trained with tlno_gpu_mlp_* (NULL turns the screen
 * off).  Following builds evaluate the network on the matrix cores for every
 * candidate shell quartet and replace Q_ab Q_cd by the estimate
 * est = Q_ab Q_cd min(1, 10^(y + margin)) in the density test, in the Schwarz
 * test and in the FP32 gate (y: predicted log10(max|(ab|cd)| / (Q_ab Q_cd)),
 * margin: safety offset in decades). */
int tlno_gpu_set_learned_screen(tlno_gpu_plan* plan, const float* params, double margin);

/* The 32 learned-screen features of n quartets (pair_p[i] | pair_q[i]) (plan
 * pair indices), out[32 i + k]. */
int tlno_gpu_quartet_features(tlno_gpu_plan* plan, int n, const int* pair_p, const int* pair_q,
                              float* out);

/* Per-quartet routing statistics of the last build (zeros if it did not use
 * the density test or the learned screen): stats[0] candidate quartets,
 * [1] kept by the density test, [2] kept after the learned screen,
 * [3] of those sent to FP32. */
int tlno_gpu_last_route_stats(const tlno_gpu_plan* plan, long long* stats);

/* Kernel policy (no reference counterpart): FP64 builds evaluate classes
 * with at least min_outputs Cartesian values per shell quartet with the team
 * evaluator (one wave per quartet, horizontal recurrence after the primitive
 * contraction); smaller classes, and every class when min_outputs = 0, use the
 * lane evaluator (one lane per quartet).  Default 0 (tlno_gpu_plan_autotune
 * chooses per class). */
int tlno_gpu_set_team_threshold(tlno_gpu_plan* plan, int min_outputs);

/* Autotuner of the plan's kernels (no reference counterpart): per class pair,
 * times the lane and the team evaluator on this plan (repeats serialized FP64
 * builds each) and keeps the faster for following builds, and times the lane
 * evaluator once more with every quartet on its FP32 kernels: a class whose
 * FP32 kernels are not faster runs its FP32 share (tlno_gpu_set_mixed_precision)
 * on the FP64 kernels.  report receives a JSON summary per class pair
 * (lane_ms, team_ms, lane_fp32_ms, choice; may be NULL).
 * tlno_gpu_set_team_threshold discards the choices. */
int tlno_gpu_plan_autotune(tlno_gpu_plan* plan, double schwarz, int repeats, char* report,
                           int report_len);

/* Mixed precision (no reference counterpart): with fp32_threshold > 0,
 * following builds of this plan evaluate the 2D recurrences and the Cartesian
 * contraction of every shell quartet with Q_ab Q_cd max|D| < fp32_threshold in
 * FP32 (Rys roots and weights, T and the prefactor stay FP64; every J/K
 * accumulation is FP64).  0 restores FP64 everywhere (the default).
 * tlno_gpu_last_mixed_stats reports the FP32 share of the last build and the
 * max|D| it gated with. */
int tlno_gpu_set_mixed_precision(tlno_gpu_plan* plan, double fp32_threshold);
int tlno_gpu_last_mixed_stats(const tlno_gpu_plan* plan, long long* fp32_shell_quartets,
                              long long* fp32_prim_quartets, double* max_abs_density);

/* Work statistics of the last build: shell quartets and primitive quartets
 * evaluated (after Schwarz screening). */
int tlno_gpu_last_build_stats(const tlno_gpu_plan* plan, long long* shell_quartets,
                              long long* prim_quartets);

/* Per-class-pair kernel timing (no reference counterpart; measurement hook):
 * enable = 1 records HIP events on the plan's stream around the kernels of
 * each bra/ket class pair of every following build and resets the
 * accumulators.  tlno_gpu_class_stats returns, for the last build's
 * threshold, 36 entries each (index 6*X + Y, classes ss ps pp ds dp dd):
 * shell quartets, primitive quartets and accumulated milliseconds over
 * `builds` timed builds. */
int tlno_gpu_set_timing(tlno_gpu_plan* plan, int enable);
int tlno_gpu_class_stats(const tlno_gpu_plan* plan, long long* shell_quartets,
                         long long* prim_quartets, double* ms, long long* builds);
/* The same over all fifteen pair classes (ss ps pp ds dp dd fs fp fd ff gs
 * gp gd gf gg): 225 entries each, index 15*X + Y for bra class X >= ket class
 * Y (the f and g class pairs time their team-evaluator launch; other
 * entries 0). */
int tlno_gpu_class_stats_all(const tlno_gpu_plan* plan, long long* shell_quartets,
                             long long* prim_quartets, double* ms, long long* builds);

/* Unit-diagonal Cartesian ERI blocks (ga gb|gc gd) for n group quadruples
 * (groups4[4i..4i+3], groups numbered in first-appearance order); block i is
 * written at out + offsets[i] as [a][b][c][d] in each group's component
 * order (the order the components appear in the basis).  Test/audit entry
 * point (audit-TU counterpart: compute_quartet_block_rys_impl,
 * rys_engine.ipp:13754-13842). */
int tlno_gpu_eri_cart_blocks(tlno_gpu_plan* plan, int n, const int* groups4,
                             const long long* offsets, double* out);

/* Boys F_0..F_mmax at n arguments: out[i*(mmax+1)+m], 0 <= mmax <=
 * TLNO_GPU_BOYS_MMAX.  Counterpart of boys_single / boys_r000_all
 * (core/math/general_math.hpp:6-8): orders up to 16 use the device evaluator
 * of the ERI kernels, higher orders a series plus downward recursion.  Every
 * T must be finite and >= 0 (checked; an error otherwise). */
#define TLNO_GPU_BOYS_MMAX 128
int tlno_gpu_boys(int n, const double* T, int mmax, double* out);

/* Rys roots t^2 and weights for N = 1..9 at n arguments: roots[i*N+k]; every
 * T finite and >= 0 (checked).
 * Counterpart of rys_roots_weights (rys_engine.ipp:3820-3837). */
int tlno_gpu_rys(int N, int n, const double* T, double* roots, double* weights);

/* One-electron S, T, V (n x n row-major each, n = M of the plan's basis in
 * the plan's mode).  Replaces build_one_electron_matrices /
 * build_one_electron_spherical_matrices (integrals/one_electron/onee_matrices.hpp:31-79).
 * nuclear charges Z (n_nuc) and centres R (3*n_nuc, bohr). */
int tlno_gpu_onee(tlno_gpu_plan* plan, int n_nuc, const double* Z, const double* R,
                  double* S, double* T, double* V);

/* The same S, T, V with the reference's one-electron screening
 * (integrals/one_electron/onee_matrices.cpp:100-210 group pairs,
 * onee_integrals.cpp:541-556 nuclear primitive pairs): shell pairs whose decay
 * estimate K_a K_b exp(-mu_min AB^2) is below decay_eps are dropped, kept
 * pairs whose overlap bound is below it skip V, and the V contribution of a
 * primitive pair whose monopole estimate is below nuclear_eps / sum|Z| is
 * skipped (0 turns either screen off; the reference's defaults are 1e-15 and
 * 1e-15).  tlno_gpu_onee is this call with both thresholds 0.  Any output may
 * be NULL.  stats (6 doubles, may be NULL): shell groups, possible group
 * pairs G(G+1)/2, kept pairs, primitive pairs of the pairs that evaluate V,
 * of those kept by the nuclear screen (the counts of the reference's
 * tlno_onee_pair_stats / tlno_onee_nuclear_pair_stats,
 * benchmarks/one_electron_engine/src/onee_capi.cpp:168-209), and the device
 * milliseconds of the S/T/V kernels. */
int tlno_gpu_onee_screened(tlno_gpu_plan* plan, int n_nuc, const double* Z, const double* R,
                           double decay_eps, double nuclear_eps, double* S, double* T, double* V,
                           double* stats);

/* ---- molecule batches (no reference counterpart: the reference builds one
 * molecule per call).  n_mol molecules in the Cartesian component basis,
 * their component arrays concatenated in molecule order (n_comp[m]
 * components each, same per-component layout as tlno_gpu_plan_create).
 * D, J and K of molecule m are M_m x M_m blocks packed at offsets[m] of one
 * array of `total` doubles (tlno_gpu_batch_info).  Every class kernel runs
 * once for all molecules, persistent with the launch shape of its class
 * (block size, waves per CU) chosen by tlno_gpu_batch_autotune or set by
 * tlno_gpu_batch_set_shape (default: 256 threads, one wave per task).
 * use_graph = 1 captures the whole build into a HIP graph on the first call
 * for a given (density, J, K, schwarz) and replays it afterwards. */
typedef struct tlno_gpu_batch tlno_gpu_batch;
int tlno_gpu_batch_create(int n_mol, const int* n_comp, const double* centers, const int* ang,
                          const int* nprim, const double* exps, const double* coefs,
                          tlno_gpu_batch** out);
void tlno_gpu_batch_destroy(tlno_gpu_batch* batch);
int tlno_gpu_batch_info(const tlno_gpu_batch* batch, int* n_mol, long long* total, int* M,
                        long long* offsets);
int tlno_gpu_batch_build_jk(tlno_gpu_batch* batch, const double* density, double* J, double* K,
                            double schwarz);
int tlno_gpu_batch_build_jk_device(tlno_gpu_batch* batch, const double* d_density, double* d_J,
                                   double* d_K, double schwarz, int use_graph);
/* Times every candidate launch shape per class pair (repeats launches each)
 * and keeps the fastest; report receives a JSON summary (may be NULL). */
int tlno_gpu_batch_autotune(tlno_gpu_batch* batch, double schwarz, int repeats, char* report,
                            int report_len);
int tlno_gpu_batch_set_shape(tlno_gpu_batch* batch, int bra_class, int ket_class, int block,
                             int waves_per_cu);
/* Shell and primitive quartets and device milliseconds of the last build. */
int tlno_gpu_batch_last_stats(const tlno_gpu_batch* batch, long long* shell_quartets,
                              long long* prim_quartets, double* ms);

/* ---- learned-screen network: 32 -> 64 -> 64 -> 1, ReLU, bf16 weights,
 * FP32 sums, trained on the GPU (no reference counterpart).  Parameters are
 * TLNO_GPU_MLP_NALL floats: W1[64][32], b1[64], W2[64][64], b2[64], w3[64],
 * b3, feature mean[32], feature 1/std[32]. */
#define TLNO_GPU_MLP_NFEAT 32
#define TLNO_GPU_MLP_NPARAM 6337
#define TLNO_GPU_MLP_NALL 6401
typedef struct tlno_gpu_mlp tlno_gpu_mlp;
/* He-uniform initialisation from seed, identity feature normalization. */
int tlno_gpu_mlp_create(int device, unsigned seed, tlno_gpu_mlp** out);
void tlno_gpu_mlp_destroy(tlno_gpu_mlp* mlp);
int tlno_gpu_mlp_get_params(tlno_gpu_mlp* mlp, float* params);
/* Replaces all parameters and resets the optimizer state. */
int tlno_gpu_mlp_set_params(tlno_gpu_mlp* mlp, const float* params);
/* Sets the feature normalization to the mean and 1/std of n samples X[32 i + k]. */
int tlno_gpu_mlp_set_normalization(tlno_gpu_mlp* mlp, long long n, const float* X);
/* Loss mean_i c(r_i) r_i^2 (r = t - y, c = w_under if r > 0 else 1) over n
 * samples and its gradient with respect to the TLNO_GPU_MLP_NPARAM parameters. */
int tlno_gpu_mlp_loss_grad(tlno_gpu_mlp* mlp, long long n, const float* X, const float* t,
                           float w_under, float* grad, double* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the full batch. */
int tlno_gpu_mlp_adam_step(tlno_gpu_mlp* mlp, long long n, const float* X, const float* t,
                           float lr, float w_under);
/* Mini-batch Adam over shuffled epochs, cosine decay from lr to lr/20;
 * epoch_loss[e] = mean batch loss of epoch e (may be NULL). */
int tlno_gpu_mlp_train(tlno_gpu_mlp* mlp, long long n, const float* X, const float* t, int epochs,
                       int batch, float lr, float w_under, unsigned seed, float* epoch_loss);
/* Network output for n samples; matrix_cores = 1 uses the bf16 MFMA evaluator
 * of the screening kernel, 0 the per-lane FP32 evaluator. */
int tlno_gpu_mlp_predict(tlno_gpu_mlp* mlp, long long n, const float* X, float* y, int matrix_cores);

#ifdef __cplusplus
}
#endif
#endif
