/*
 * tlno_onee.h — the reference's one-electron C ABI, re-exported by the C++
 * facade library (origo-tlno_amd/tlno_amd/libtlno_amd_cpp.so) over the GPU
 * engine.  Same names, signatures and behaviour as the reference's
 * benchmarks/one_electron_engine/src/onee_capi.cpp:36-225 and
 * basis_export.cpp:43: raw pointers, row-major n x n doubles, caller-owned
 * outputs; functions never throw; a null handle makes every call a no-op
 * (queries return -1).  S/T/V are evaluated on the GPU with the reference's
 * screening rules (group-pair decay estimate fixed at plan build, nuclear
 * primitive-pair estimate at evaluation).
 *
 * Basis input differs in one respect: the reference's named basis tables are
 * not part of this build, so basis_name is the path of a JSON basis file
 * ("*.json") holding, per element symbol, the object tlno_onee_basis_json
 * returns: {"charge": Z, "shells": [[L, [exponent, coefficient], ...], ...]}.
 * A named basis ("sto-3g", "def2-svp", ...) returns nullptr, as the reference
 * does for names it does not know.
 */
#ifndef TLNO_ONEE_H
#define TLNO_ONEE_H

#ifdef __cplusplus
extern "C" {
#endif

/* onee_capi.cpp:36-58: molecule (".inp", "* xyz q m" block, bohr) + basis;
 * nullptr on any failure. */
void* tlno_onee_create_from_inp(const char* inp_path, const char* basis_name);

/* No reference counterpart: the same handle from a component basis in the
 * layout of tlno_gpu_plan_create (normalized coefficients) and nuclei
 * (integer charges, 3 n_nuc bohr coordinates); nullptr on failure. */
void* tlno_onee_create(int n_comp, const double* centers, const int* ang, const int* nprim,
                       const double* exps, const double* coefs, int n_nuc, const int* charges,
                       const double* nuc_centers);

/* onee_capi.cpp:60-75: Cartesian and spherical AO counts (nbf_sph builds the
 * plan when there is none). */
int tlno_onee_nbf(void* handle);
int tlno_onee_nbf_sph(void* handle);

/* onee_capi.cpp:77-105: global screening switches (screening on resets the
 * decay and nuclear thresholds to 1e-15, off sets them to 0). */
void tlno_onee_set_screening(int enabled);
int tlno_onee_screening_enabled(void);
void tlno_onee_set_group_pair_decay_eps(double eps);
double tlno_onee_group_pair_decay_eps(void);
void tlno_onee_set_nuclear_screen_eps(double eps);
double tlno_onee_nuclear_screen_eps(void);

/* onee_capi.cpp:107-146: plan once (prepare_plan) and evaluate with it
 * (eval_cached[_sph]), or rebuild the plan and evaluate (eval_full[_sph]). */
void tlno_onee_prepare_plan(void* handle);
void tlno_onee_eval_full(void* handle);
void tlno_onee_eval_cached(void* handle);
void tlno_onee_eval_full_sph(void* handle);
void tlno_onee_eval_cached_sph(void* handle);

/* onee_capi.cpp:148-166: the last Cartesian / spherical S, T, V (any NULL
 * pointer is skipped). */
void tlno_onee_copy(void* handle, double* S, double* T, double* V);
void tlno_onee_copy_sph(void* handle, double* S, double* T, double* V);

/* onee_capi.cpp:168-209: shell groups, possible and kept group pairs of the
 * plan; primitive pairs of the pairs that evaluate V, and of those kept by
 * the nuclear screen at the current threshold. */
void tlno_onee_pair_stats(void* handle, int* n_groups, int* total_possible, int* kept);
void tlno_onee_nuclear_pair_stats(void* handle, long long* total, long long* kept);

/* onee_capi.cpp:212-219 */
void tlno_onee_noop(void* handle);
void tlno_onee_destroy(void* handle);

/* basis_export.cpp:43-89: one element's basis parameters as JSON (thread-
 * local string valid until the next call on the thread); nullptr when
 * unsupported.  Here basis_name is a JSON basis file. */
const char* tlno_onee_basis_json(const char* basis_name, const char* symbol);

#ifdef __cplusplus
}
#endif
#endif
