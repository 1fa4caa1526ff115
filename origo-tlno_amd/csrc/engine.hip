// Host engine and C ABI: plan construction (shell grouping, primitive-pair
// SoA, GPU Schwarz bounds, class/bound ordering, work lists), density and
// result transforms, and the J/K build driver.
//
// Plan construction follows the reference's plan semantics
// (integrals/two_electron/rys/rys_engine.ipp:1966-2062 grouping,
// 2243-2277 primitive pairs, 2365-2460 AO numbering and pair orientation,
// 16164-16258 Schwarz bounds) but lays everything out for the GPU:
//   * internal Cartesian AO order = shells in first-appearance order, each
//     shell's components contiguous in descending-lex order, so a shell's
//     block is [base, base + ncart) in every M x M matrix;
//   * shell pairs sorted by class (ss, ps, pp, ds, dp, dd) and, inside a
//     class, by descending Schwarz bound, so the surviving kets of a bra are
//     a prefix found by binary search;
//   * primitive pairs as {Px, Py, Pz, p} double4 + weight arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"
#include "eri_device.hpp"
#include "mlp.hpp"
#include "onee.hpp"
#include "tlno_gpu.h"

// Classes with at least this many Cartesian values per shell quartet run the
// team evaluator (one wave per quartet, HRR after contraction) in FP64 builds;
// 0 (default): the lane evaluator unless the plan's autotuner chose the team
// evaluator for a class.
#ifndef TLNO_TEAM_MIN_OUT
#define TLNO_TEAM_MIN_OUT 0
#endif

namespace tlno_amd {

thread_local std::string g_err;

// ------------------------------------------------------------- device ctx --
// Class-pair kernels of one build run concurrently on up to this many
// streams, shared by the plans of a device (created on the first build).
// HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4
// unless the process environment sets the variable before the HIP runtime
// loads); min(that, kSideStreams) streams are used, since more streams than
// queues only serialize behind each other.
#ifndef tlno_amd_side_streams
#define tlno_amd_side_streams 8
#endif
constexpr int kSideStreams = tlno_amd_side_streams;

// total angular momentum and Cartesian components of the pair classes ss,
// ps, pp, ds, dp, dd (lane and team evaluators) and fs, fp, fd, ff, gs, gp,
// gd, gf, gg (team evaluator only, their pairs after all others in plan order)
constexpr int kClassLAB[kNumClasses] = {0, 1, 2, 2, 3, 4, 3, 4, 5, 6, 4, 5, 6, 7, 8};
constexpr int kClassNcart[kNumClasses] = {1, 3, 9, 6, 18, 36, 10, 30, 60, 100, 15, 45, 90, 150, 225};
constexpr int kClassL[kNumClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}, {3, 0}, {3, 1},
                                         {3, 2}, {3, 3}, {4, 0}, {4, 1}, {4, 2}, {4, 3}, {4, 4}};
constexpr int kNumHigh = kNumClasses - kNumLowClasses;
constexpr int kMaxComp = ncart(kMaxShellL);  // Cartesian components of the largest shell

int side_streams_for_queues() {
    int q = 4;  // HIP's default
    if (const char* v = std::getenv("GPU_MAX_HW_QUEUES")) {
        const int n = std::atoi(v);
        if (n > 0) q = n;
    }
    return std::max(1, std::min(kSideStreams, q));
}

// ---- allocation-free builds ----
// Scratch of tlno_gpu_build_jk_device_noalloc: the internal-order density and
// the J/K accumulators (3 x Mcart^2 doubles, Mcart <= kNoAllocMaxM) followed
// by the work counters (4 x kCounterSlots) and one scalar block.  A static
// device array: every GPU the library runs on holds its own copy (24 MiB +
// 4 KiB) until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns per device (DeviceCtx::scratch).
constexpr int kNoAllocMaxM = 1024;
constexpr std::size_t kScratchTail = 4 * kCounterSlots + 8;
constexpr std::size_t kScratchDoubles = 3 * static_cast<std::size_t>(kNoAllocMaxM) * kNoAllocMaxM + kScratchTail;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Set while an allocation-free build runs on this thread: every lazily
// allocating step (work lists of a new threshold, routing buffers, team
// quartet lists, side streams) raises instead of allocating.
thread_local bool g_no_alloc = false;

void require_alloc(const char* what) {
    if (g_no_alloc)
        throw std::runtime_error(std::string("tlno_gpu_build_jk_device_noalloc: ") + what +
                                 " not prepared; run tlno_gpu_build_jk_device once on this plan with the same "
                                 "threshold and settings first");
}

struct DeviceCtx {
    bool ready = false;
    double* scratch = nullptr;  // this device's copy of the static scratch array
    double* boys = nullptr;
    double* boys_lds = nullptr;
    double* rys = nullptr;
    double* c2s = nullptr;  // spherical projection table (C2S::c)
    DevTables tb{};
    // side streams of concurrent builds; a build holds side_mu from its fork
    // to its join, so the fork/join events are never re-recorded mid-build
    std::mutex side_mu;
    int nside = 0;  // 0: not created yet
    hipStream_t side[kSideStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSideStreams] = {};
    // staging buffer of the small utility entry points (Boys, Rys), grown on
    // demand and kept for the process; used under util_mu
    std::mutex util_mu;
    double* util = nullptr;
    std::size_t util_n = 0;
    double* util_buffer(std::size_t n) {
        if (n > util_n) {
            if (util) HIP_OK(hipFree(util));
            util = nullptr;
            util_n = 0;
            HIP_OK(hipMalloc(&util, n * sizeof(double)));
            util_n = n;
        }
        return util;
    }
};

// creates the device's side streams and events (caller holds side_mu)
void ensure_side_streams(DeviceCtx& c) {
    if (c.nside > 0) return;
    require_alloc("side streams of this device");
    const int n = side_streams_for_queues();
    for (int s = 0; s < n; ++s) {
        HIP_OK(hipStreamCreateWithFlags(&c.side[s], hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&c.ev_join[s], hipEventDisableTiming));
    }
    HIP_OK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    c.nside = n;
}

// Cartesian -> spherical projection of each shell L (rows s, columns the
// local Cartesian components), built by c2s_table() below
struct C2S {
    double c[kMaxShellL + 1][2 * kMaxShellL + 1][kMaxComp] = {};
};
const C2S& c2s_table();

ClassKernels g_kernels[6][6];
HighKernels g_high[kNumHigh][kNumClasses];
std::once_flag g_kernels_once;
std::mutex g_dev_mutex;
DeviceCtx g_dev[16];
thread_local int g_device = 0;

void register_all_kernels() {
    register_kernels_bra0(g_kernels);
    register_kernels_bra1(g_kernels);
    register_kernels_bra2(g_kernels);
    register_kernels_bra3(g_kernels);
    register_kernels_bra4(g_kernels);
    register_kernels_bra4b(g_kernels);
    register_kernels_bra5(g_kernels);
    register_kernels_bra5b(g_kernels);
    register_kernels_bra5c(g_kernels);
    register_kernels_hi6(g_high);
    register_kernels_hi7(g_high);
    register_kernels_hi8a(g_high);
    register_kernels_hi8b(g_high);
    register_kernels_hi9a(g_high);
    register_kernels_hi9b(g_high);
    register_kernels_hi9c(g_high);
    register_kernels_hi10(g_high);
    register_kernels_hi11a(g_high);
    register_kernels_hi11b(g_high);
    register_kernels_hi12a(g_high);
    register_kernels_hi12b(g_high);
    register_kernels_hi12c(g_high);
    register_kernels_hi13a(g_high);
    register_kernels_hi13b(g_high);
    register_kernels_hi13c(g_high);
    register_kernels_hi13d(g_high);
    register_kernels_hi14a(g_high);
    register_kernels_hi14b(g_high);
    register_kernels_hi14c(g_high);
    register_kernels_hi14d(g_high);
    register_kernels_hi14e(g_high);
    register_kernels_hi14f(g_high);
}

DeviceCtx& device_ctx(int dev) {
    std::call_once(g_kernels_once, register_all_kernels);
    if (dev < 0 || dev >= 16) throw std::runtime_error("tlno_gpu: device index out of range");
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    DeviceCtx& c = g_dev[dev];
    if (c.ready) return c;
    HIP_OK(hipSetDevice(dev));
    const HostTables& t = host_tables();
    HIP_OK(hipMalloc(&c.boys, t.boys.size() * sizeof(double)));
    HIP_OK(hipMemcpy(c.boys, t.boys.data(), t.boys.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&c.boys_lds, t.boys_lds.size() * sizeof(double)));
    HIP_OK(hipMemcpy(c.boys_lds, t.boys_lds.data(), t.boys_lds.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&c.rys, t.rys.size() * sizeof(double)));
    HIP_OK(hipMemcpy(c.rys, t.rys.data(), t.rys.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipGetSymbolAddress(reinterpret_cast<void**>(&c.scratch), HIP_SYMBOL(scratch)));
    HIP_OK(hipMalloc(&c.c2s, sizeof(C2S::c)));
    HIP_OK(hipMemcpy(c.c2s, c2s_table().c, sizeof(C2S::c), hipMemcpyHostToDevice));
    c.tb.boys = c.boys;
    c.tb.boys_lds = c.boys_lds;
    c.tb.rys = c.rys;
    for (int i = 0; i < RYS_NMAX + 2; ++i) c.tb.rys_off[i] = t.rys_off[i];
    std::memcpy(c.tb.herm_r2, t.herm_r2, sizeof(c.tb.herm_r2));
    std::memcpy(c.tb.herm_w, t.herm_w, sizeof(c.tb.herm_w));
    c.ready = true;
    return c;
}

// -------------------------------------------------- cartesian -> spherical --
// PySCF-ordered real solid harmonics over unit-diagonal Cartesians, from the
// solid-harmonic recursion (Helgaker/Jorgensen/Olsen eqs. 6.4.70-6.4.73),
// normalized in the unit-diagonal metric.  Row s, column = local Cartesian
// component (descending lex).

double dfact(int n) {
    double v = 1.0;
    for (int k = 2 * n - 1; k > 0; k -= 2) v *= k;
    return v;
}

int local_index(int L, int lx, int ly) {
    const int dx = L - lx;
    return dx * (dx + 1) / 2 + (dx - ly);
}

void local_tuple(int L, int idx, int* t) {
    int k = 0;
    for (int x = L; x >= 0; --x)
        for (int y = L - x; y >= 0; --y, ++k)
            if (k == idx) {
                t[0] = x;
                t[1] = y;
                t[2] = L - x - y;
                return;
            }
}

const C2S& c2s_table() {
    static C2S tab = [] {
        C2S out;
        // polynomials over monomials of degree l, index local_index
        std::vector<std::vector<std::vector<double>>> S(kMaxShellL + 1);
        for (int l = 0; l <= kMaxShellL; ++l)
            S[l].assign(2 * l + 1, std::vector<double>(static_cast<std::size_t>(ncart(l)), 0.0));
        S[0][0][0] = 1.0;
        auto mul_axis = [](const std::vector<double>& a, int L, int axis, double s, std::vector<double>& o) {
            for (int i = 0; i < ncart(L); ++i) {
                if (a[i] == 0.0) continue;
                int t[3];
                local_tuple(L, i, t);
                t[axis] += 1;
                o[local_index(L + 1, t[0], t[1])] += s * a[i];
            }
        };
        auto mul_r2 = [](const std::vector<double>& a, int L, double s, std::vector<double>& o) {
            for (int i = 0; i < ncart(L); ++i) {
                if (a[i] == 0.0) continue;
                int t[3];
                local_tuple(L, i, t);
                for (int k = 0; k < 3; ++k) {
                    int u[3] = {t[0], t[1], t[2]};
                    u[k] += 2;
                    o[local_index(L + 2, u[0], u[1])] += s * a[i];
                }
            }
        };
        for (int l = 0; l < kMaxShellL; ++l) {
            const double f = std::sqrt((l == 0 ? 2.0 : 1.0) * (2.0 * l + 1.0) / (2.0 * l + 2.0));
            mul_axis(S[l][2 * l], l, 0, f, S[l + 1][2 * l + 2]);
            if (l > 0) mul_axis(S[l][0], l, 1, -f, S[l + 1][2 * l + 2]);
            mul_axis(S[l][2 * l], l, 1, f, S[l + 1][0]);
            if (l > 0) mul_axis(S[l][0], l, 0, f, S[l + 1][0]);
            for (int m = -l; m <= l; ++m) {
                const double den = std::sqrt(double(l + m + 1) * (l - m + 1));
                mul_axis(S[l][m + l], l, 2, (2.0 * l + 1.0) / den, S[l + 1][m + l + 1]);
                if (l > 0 && std::abs(m) <= l - 1)
                    mul_r2(S[l - 1][m + l - 1], l - 1, -std::sqrt(double(l + m) * (l - m)) / den,
                           S[l + 1][m + l + 1]);
            }
        }
        for (int L = 0; L <= kMaxShellL; ++L) {
            const int nc = ncart(L);
            for (int s = 0; s < 2 * L + 1; ++s) {
                const int m = (L == 1) ? (s == 0 ? 1 : (s == 1 ? -1 : 0)) : s - L;
                const auto& P = S[L][m + L];
                double row[kMaxComp];
                for (int c = 0; c < nc; ++c) {
                    int t[3];
                    local_tuple(L, c, t);
                    row[c] = P[c] * std::sqrt(dfact(t[0]) * dfact(t[1]) * dfact(t[2]));
                }
                double n2 = 0.0;
                for (int c = 0; c < nc; ++c)
                    for (int d = 0; d < nc; ++d) {
                        int tc[3], td[3];
                        local_tuple(L, c, tc);
                        local_tuple(L, d, td);
                        double g = 1.0;
                        for (int k = 0; k < 3; ++k) {
                            const int sm = tc[k] + td[k];
                            if (sm % 2) {
                                g = 0.0;
                                break;
                            }
                            g *= dfact(sm / 2) / std::sqrt(dfact(tc[k]) * dfact(td[k]));
                        }
                        n2 += row[c] * row[d] * g;
                    }
                for (int c = 0; c < nc; ++c) out.c[L][s][c] = row[c] / std::sqrt(n2);
            }
        }
        return out;
    }();
    return tab;
}

// ------------------------------------------------------------------ plan --
struct GroupH {
    int L = 0;
    double C[3] = {0, 0, 0};
    int nprim = 0;
    std::vector<double> exps, cr;  // radial coefficients c / N_A
    int comp[15];                  // basis index of each local canonical component
    int first = 0;                 // first-appearance component
    double na_first = 1.0;         // angular normalization of that component
    int int_base = 0;
    int sph_base = 0;
};


struct TaskSet {
    int4* d_tasks = nullptr;
    long long off[6][6] = {};
    long long cnt[6][6] = {};
    long long shell_quartets = 0;
    long long prim_quartets = 0;
    long long cls_sq[6][6] = {};  // shell quartets per class pair
    long long cls_pq[6][6] = {};  // primitive quartets per class pair
    // per class pair, in task order: max bound of each task (descending) and
    // running shell/primitive quartet counts (for the precision split)
    std::vector<double> bmax[6][6];
    std::vector<long long> cum_sq[6][6], cum_pq[6][6];
    // Bucketed list tasks, for ket classes whose primitive-pair counts differ
    // (a wave takes as long as its heaviest ket): d_bk holds, per ket class,
    // its pairs grouped by bucket of primitive-pair count (each group in
    // bound order), and list tasks {P, offset into d_bk, count} hold at most
    // 64 kets of one group; lcnt = 0 where the range tasks are used.
    int* d_bk = nullptr;
    int4* d_ltasks = nullptr;
    int4* d_ltasks_lpt = nullptr;  // the same list tasks by descending cost (inside d_ltasks' allocation)
    long long loff[6][6] = {}, lcnt[6][6] = {};
    std::vector<double> lbmax[6][6];
    // per-quartet routing (density test, learned screen): rows of at most
    // kRouteSeg candidate kets {P, k_begin, k_end, 6 X + Y}, the qlist region
    // of each row, and the task capacity of each class pair; built on first use
    int4* d_rows = nullptr;
    long long* d_rbase = nullptr;
    long long* d_rtask = nullptr;  // first task slot of each row
    int* d_rcnt = nullptr;         // learned builds: density survivors per row
    long long nrows = 0;
    long long qtotal = 0;
    long long toff[36] = {}, tcap[36] = {};
    long long tcap_total = 0;
    // team evaluator work lists: per class pair, the shell quartets {P, Q}
    // of its range tasks in task order (cum_sq gives the prefix of the first
    // n tasks); built on the first team launch of the class
    int2* d_tq[6][6] = {};
    // quartets with an f shell (team evaluator only): per class pair
    // (X >= 6, Y <= X) the Schwarz-surviving {P, Q}, heaviest first
    int2* d_hq = nullptr;
    long long hoff[kNumHigh][kNumClasses] = {}, hcnt[kNumHigh][kNumClasses] = {};
    long long hpq[kNumHigh][kNumClasses] = {};  // their primitive quartets
};

}  // namespace tlno_amd

struct tlno_gpu_plan {
    int device = 0;
    int spherical = 0;
    int Mcart = 0, Msph = 0, Mout = 0;
    std::vector<tlno_amd::GroupH> groups;
    std::vector<int> int2basis;
    int npair = 0;
    int cls_begin[tlno_amd::kNumClasses + 1] = {0};
    std::vector<int> pgA, pgB;
    std::vector<double> Q;
    long long nprim_pairs = 0;
    std::vector<int> pn_host;
    // host mirror of the pair data in plan order
    std::vector<double> h_A, h_AB, h_pw;
    std::vector<int> h_ao, h_poff, h_diag;
    std::vector<double4> h_pP;
    // device pair data
    double *dA = nullptr, *dAB = nullptr, *dpw = nullptr;
    int *dao = nullptr, *dpoff = nullptr, *dpn = nullptr, *ddiag = nullptr;
    double4* dpP = nullptr;
    double* dpexp = nullptr;  // [nprim][2] alpha, beta (one-electron kinetic term)
    int* dpL = nullptr;       // [npair][2] LA, LB
    // transform metadata (per internal Cartesian index / per output index)
    int *d_int2basis = nullptr, *d_intL = nullptr, *d_intloc = nullptr, *d_intsb = nullptr;
    int *d_sphL = nullptr, *d_sphloc = nullptr, *d_sphib = nullptr;
    // work buffers
    double *dDin = nullptr, *dDint = nullptr, *dJacc = nullptr, *dKacc = nullptr;
    double *dJout = nullptr, *dKout = nullptr;
    hipStream_t stream = nullptr;
    std::map<double, tlno_amd::TaskSet> tasks;
    long long last_shell_quartets = 0, last_prim_quartets = 0;
    double last_thr = -1.0;
    // mixed precision: shell quartets with Q_ab Q_cd max|D| < fp32_gate run
    // the FP32 recurrences (0 = FP64 everywhere)
    double fp32_gate = 0.0;
    double* d_scalar = nullptr;  // device scratch for max|D|
    // density screen (the reference's density_threshold): 0 = off
    double density_thr = 0.0;
    // heuristic screens of the reference's ScreeningOptions: primitive
    // FASTSCHWARZ threshold and shell-pair decay indicator threshold (0: off),
    // decay indicator per pair (plan order)
    double fastschwarz = 0.0, decay_thr = 0.0;
    double* d_decay = nullptr;
    int* d_row_group = nullptr;              // [Mout] shell of each row of the caller's basis
    int *d_pgA = nullptr, *d_pgB = nullptr, *d_pcls = nullptr, *d_clsb = nullptr;
    double* d_Q = nullptr;                   // [npair] bounds in plan order
    unsigned long long* d_gmax = nullptr;    // [ngroups]
    unsigned long long* d_clsmax = nullptr;  // [kNumClasses]
    double* d_pairfac = nullptr;             // [npair]
    int* d_cut = nullptr;                    // [6 npair]
    // global scratch of the g classes whose HRR and digestion outgrow LDS:
    // gslots workgroup slots of the largest such class's working set
    double* d_gwork = nullptr;
    long long gwork_doubles = 0;
    int gslots = 0;
    unsigned long long* d_counter = nullptr; // shell quartets evaluated by the last build
    long long last_evaluated = 0, last_evaluated_prim = 0;
    double last_maxD = 0.0;
    long long last_fp32_sq = 0, last_fp32_pq = 0;
    // per-quartet routing: pair features, shell rows, shell-block max |D|,
    // learned-screen operands, and the work lists the routing kernel fills
    float4 *d_pf0 = nullptr, *d_pf1 = nullptr, *d_pcen = nullptr;
    int* d_grows = nullptr;                  // [ngroups][kMaxComp] rows of each shell in the caller's basis (-1 pad)
    double* d_blk = nullptr;                 // [ngroups^2] max |D| over each shell block
    double* d_bsum = nullptr;                // [ngroups^2] sum |D| over each shell block
    int learned = 0;
    double learned_margin = 0.0;
    uint4* d_frag = nullptr;
    float *d_mvec = nullptr, *d_mnorm = nullptr;
    float mb3 = 0.0f;
    int* d_qlist = nullptr;
    long long qlist_cap = 0;
    int* d_cand = nullptr;  // learned builds: density survivors, laid out like qlist
    long long cand_cap = 0;
    int4 *d_t64 = nullptr, *d_t32 = nullptr;
    long long tlist_cap = 0;
    unsigned long long* d_rstats = nullptr;  // [kCounterSlots][4]
    long long last_route[4] = {0, 0, 0, 0};  // candidates, density-kept, kept, FP32 (last routed build)
    int team_min_out = TLNO_TEAM_MIN_OUT;    // 0: lane evaluator for every class
    signed char team_sel[6][6];              // autotuned choice per class pair: -1 none, 0 lane, 1 team
    double cls_cost[6][6] = {};              // measured ms of the chosen evaluator per class pair (0: not measured)
    // autotuned: the FP32 share of a lane class runs the FP32 kernels only
    // where they measured faster than the FP64 ones (an FP32 FMA issues at
    // the FP64 rate without packed math, so the gain is occupancy, if any)
    bool fp32_slower[6][6] = {};
    // optional per-class kernel timing with HIP events on the plan stream
    int timing = 0;
    hipEvent_t ev[6][6][2] = {};
    // one-electron evaluation: kept pair list of the last decay threshold and
    // work buffers, kept between calls
    double onee_decay = -1.0;
    std::vector<int2> onee_list;
    int onee_cb[tlno_amd::kNumClasses + 1] = {0};
    long long onee_kept = 0;
    int2* d_onee_list = nullptr;
    double* d_onee_X = nullptr;  // 3 Mcart^2 internal-order S, T, V
    double* d_onee_O = nullptr;  // 3 Mout^2 caller-order S, T, V
    double* d_onee_nuc = nullptr;
    int onee_nuc_cap = 0;
    hipEvent_t onee_ev[2] = {nullptr, nullptr};
    double cls_ms[6][6] = {};
    long long timed_builds = 0;
    // timing mode: the class pairs with an f shell
    hipEvent_t hev[tlno_amd::kNumClasses - tlno_amd::kNumLowClasses][tlno_amd::kNumClasses][2] = {};
    double hcls_ms[tlno_amd::kNumClasses - tlno_amd::kNumLowClasses][tlno_amd::kNumClasses] = {};

    tlno_amd::PairData pair_data() const {
        tlno_amd::PairData p;
        p.A = dA;
        p.AB = dAB;
        p.ao = dao;
        p.poff = dpoff;
        p.pn = dpn;
        p.pP = dpP;
        p.pw = dpw;
        p.diag = ddiag;
        return p;
    }
    ~tlno_gpu_plan();
};

namespace tlno_amd {

template <class T>
void dev_upload(T*& dst, const std::vector<T>& src) {
    if (dst) HIP_OK(hipFree(dst));
    dst = nullptr;
    if (src.empty()) return;
    HIP_OK(hipMalloc(&dst, src.size() * sizeof(T)));
    HIP_OK(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
}

bool finite(double v) { return std::isfinite(v); }

int class_of(int LA, int LB) {
    // LA >= LB
    static const int t[5][5] = {{0, -1, -1, -1, -1}, {1, 2, -1, -1, -1}, {3, 4, 5, -1, -1}, {6, 7, 8, 9, -1},
                                {10, 11, 12, 13, 14}};
    return t[LA][LB];
}

// Group components as the reference does (rys_engine.ipp:1925-2062).
std::vector<GroupH> build_groups(int n, const double* centers, const int* ang, const int* nprim,
                                 const double* exps, const double* coefs) {
    std::vector<GroupH> groups;
    std::size_t off = 0;
    for (int i = 0; i < n; ++i) {
        const double* c = centers + 3 * i;
        const int* a = ang + 3 * i;
        for (int k = 0; k < 3; ++k)
            if (!finite(c[k])) throw std::runtime_error("eri_ao_tensor_hgp: shell centers must be finite.");
        if (a[0] < 0 || a[1] < 0 || a[2] < 0)
            throw std::runtime_error("eri_ao_tensor_hgp: Cartesian angular powers must be nonnegative.");
        const int np = nprim[i];
        if (np <= 0)
            throw std::runtime_error("eri_ao_tensor_hgp: contracted shells need at least one primitive.");
        if (np > 16) throw std::runtime_error("eri_ao_tensor_hgp: at most 16 primitives per shell.");
        const double* e = exps + off;
        const double* cf = coefs + off;
        off += static_cast<std::size_t>(np);
        for (int p = 0; p < np; ++p)
            if (!finite(e[p]) || e[p] <= 0.0 || !finite(cf[p]))
                throw std::runtime_error(
                    "eri_ao_tensor_hgp: primitive exponents must be finite and positive and coefficients "
                    "must be finite.");
        const int L = a[0] + a[1] + a[2];
        if (L > kMaxShellL)
            throw std::runtime_error(
                "eri_ao_tensor_hgp: this build supports shell L=0..4 (s, p, d, f, g).");
        const double na = 1.0 / std::sqrt(dfact(a[0]) * dfact(a[1]) * dfact(a[2]));
        int gi = -1;
        for (int g = 0; g < static_cast<int>(groups.size()) && gi < 0; ++g) {
            const GroupH& G = groups[g];
            if (G.L != L || G.C[0] != c[0] || G.C[1] != c[1] || G.C[2] != c[2] || G.nprim != np) continue;
            bool same = true;
            for (int p = 0; p < np && same; ++p) {
                if (G.exps[p] != e[p]) same = false;
                const double x = G.cr[p], y = cf[p] / na;
                if (x == y) continue;
                const double sc = std::max(std::abs(x), std::abs(y));
                if (sc != 0.0 && std::abs(x - y) > 32.0 * 2.220446049250313e-16 * sc) same = false;
            }
            if (same) gi = g;
        }
        if (gi < 0) {
            GroupH G;
            G.L = L;
            for (int k = 0; k < 3; ++k) G.C[k] = c[k];
            G.nprim = np;
            G.exps.assign(e, e + np);
            G.cr.resize(np);
            for (int p = 0; p < np; ++p) G.cr[p] = cf[p] / na;
            for (int k = 0; k < 15; ++k) G.comp[k] = -1;
            G.first = i;
            G.na_first = na;
            gi = static_cast<int>(groups.size());
            groups.push_back(std::move(G));
        }
        GroupH& G = groups[gi];
        const int loc = local_index(L, a[0], a[1]);
        if (G.comp[loc] >= 0)
            throw std::runtime_error(
                "eri_ao_tensor_hgp: duplicate or invalid Cartesian component inside a contracted shell.");
        G.comp[loc] = i;
    }
    for (auto& G : groups)
        for (int k = 0; k < ncart(G.L); ++k)
            if (G.comp[k] < 0)
                throw std::runtime_error(
                    "eri_ao_tensor_hgp: each radial contraction must contain exactly one complete Cartesian "
                    "shell.");
    return groups;
}

// ------------------------------------------------------ transform kernels --
// The projection table on the device ([L][s][component], the layout of
// C2S::c), passed by pointer: for shells up to g it outgrows a kernel argument.
struct C2SArgs {
    const double* c;
    __device__ double at(int L, int s, int comp) const {
        return c[(L * (2 * kMaxShellL + 1) + s) * kMaxComp + comp];
    }
};

// D_int[i][j] = D_user[perm[i]][perm[j]]  (Cartesian mode)
__global__ void k_prep_cart(const double* __restrict__ Din, double* __restrict__ Dint, const int* perm, int M) {
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)M * M) return;
    const int i = static_cast<int>(idx / M), j = static_cast<int>(idx % M);
    Dint[idx] = Din[(long)perm[i] * M + perm[j]];
}

// D_int[i][j] = sum_st C[L_i][s][loc_i] D_sph[sb_i + s][sb_j + t] C[L_j][t][loc_j]
__global__ void k_prep_sph(const double* __restrict__ Dsph, double* __restrict__ Dint, const int* iL,
                           const int* iloc, const int* isb, int Mc, int Ms, C2SArgs c2s) {
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)Mc * Mc) return;
    const int i = static_cast<int>(idx / Mc), j = static_cast<int>(idx % Mc);
    const int Li = iL[i], Lj = iL[j], li = iloc[i], lj = iloc[j];
    const double* row = Dsph + (long)isb[i] * Ms + isb[j];
    double v = 0.0;
    for (int s = 0; s < 2 * Li + 1; ++s) {
        double t2 = 0.0;
        for (int t = 0; t < 2 * Lj + 1; ++t) t2 += row[(long)s * Ms + t] * c2s.at(Lj, t, lj);
        v += c2s.at(Li, s, li) * t2;
    }
    Dint[idx] = v;
}

// X_user[perm[i]][perm[j]] = Xacc[i][j] + Xacc[j][i]
__global__ void k_fin_cart(const double* __restrict__ Xacc, double* __restrict__ Xout, const int* perm, int M) {
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)M * M) return;
    const int i = static_cast<int>(idx / M), j = static_cast<int>(idx % M);
    Xout[(long)perm[i] * M + perm[j]] = Xacc[idx] + Xacc[(long)j * M + i];
}

// X_sph[u][v] = sum_ij C[u][i] (Xacc[i][j] + Xacc[j][i]) C[v][j]
__global__ void k_fin_sph(const double* __restrict__ Xacc, double* __restrict__ Xout, const int* sL,
                          const int* sloc, const int* sib, int Mc, int Ms, C2SArgs c2s) {
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)Ms * Ms) return;
    const int u = static_cast<int>(idx / Ms), v = static_cast<int>(idx % Ms);
    const int Lu = sL[u], Lv = sL[v], su = sloc[u], sv = sloc[v];
    const int iu = sib[u], iv = sib[v];
    double acc = 0.0;
    for (int a = 0; a < ncart(Lu); ++a) {
        const double ca = c2s.at(Lu, su, a);
        if (ca == 0.0) continue;
        double t = 0.0;
        for (int b = 0; b < ncart(Lv); ++b) {
            const long ij = (long)(iu + a) * Mc + (iv + b), ji = (long)(iv + b) * Mc + (iu + a);
            t += c2s.at(Lv, sv, b) * (Xacc[ij] + Xacc[ji]);
        }
        acc += ca * t;
    }
    Xout[idx] = acc;
}

__global__ void k_fill(double* p, long n, double v) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Shell quartets of range tasks, one block of 64 threads per task: lane k
// of task i writes {P, q_begin + k} at cum[i] + k for its kets (tasks of a
// later bra primitive slice list none: slice 0 holds the task's quartets)
__global__ void k_team_quartets(const int4* __restrict__ tasks, int ntasks, const long long* __restrict__ cum,
                                int2* __restrict__ out) {
    const int i = blockIdx.x, k = threadIdx.x;
    if (i >= ntasks) return;
    const int4 t = tasks[i];
    if ((t.w & 0xffff) != 0) return;
    const int nq = min(t.z - t.y, 64);
    if (k < nq) out[cum[i] + k] = make_int2(t.x, t.y + k);
}

template <int MMAX>
__global__ void k_boys(DevTables tb, const double* T, double* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double F[MMAX + 1];
    boys_all<MMAX>(tb, T[i], F);
    for (int m = 0; m <= MMAX; ++m) out[(long)i * (MMAX + 1) + m] = F[m];
}

// F_0..F_mmax at any order (mmax above the table's orders): F_mmax from the
// positive series e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) and the
// downward recursion F_{m-1} = (2T F_m + e^{-T}) / (2m - 1), both stable;
// where T is large against mmax, the upward recursion
// F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T) from F_0 = sqrt(pi/T)/2 (erf(sqrt T)
// = 1 in binary64 for T >= 40), stable while 2T > 2m+1.
__global__ void k_boys_any(const double* T, double* out, int n, int mmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = T[i];
    double* o = out + static_cast<long>(i) * (mmax + 1);
    const double ex = exp(-t);
    if (t < 40.0 || t < 1.5 * mmax + 10.0) {
        double term = 1.0 / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 4000; ++k) {
            term *= 2.0 * t / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        double f = ex * sum;
        o[mmax] = f;
        for (int m = mmax; m >= 1; --m) {
            f = (2.0 * t * f + ex) / (2 * m - 1);
            o[m - 1] = f;
        }
    } else {
        double f = 0.88622692545275801365 / sqrt(t);
        o[0] = f;
        for (int m = 0; m < mmax; ++m) {
            f = ((2 * m + 1) * f - ex) / (2.0 * t);
            o[m + 1] = f;
        }
    }
}

template <int N>
__global__ void k_rys(DevTables tb, const double* T, double* r, double* w, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double rr[N], ww[N];
    rys_roots<N>(tb, T[i], rr, ww);
    for (int k = 0; k < N; ++k) {
        r[(long)i * N + k] = rr[k];
        w[(long)i * N + k] = ww[k];
    }
}


// grid: x covers the tasks (one wave each, 4 waves per block), y the chunks
bool use_team(const ClassKernels& k, bool fp32, const JKArgs& args, int min_out) {
    return !fp32 && k.team && !args.cw && min_out > 0 && k.nout >= min_out;
}

// the team threshold to apply to class pair (X, Y): the autotuned choice when
// there is one (1 = team, expressed as threshold 1; 0 = lane, threshold 0)
inline int class_team_min(const tlno_gpu_plan* pl, int X, int Y);

// Largest grid (in waves) of one class launch.  HIP limits a launch to
// 2^32 - 1 threads; larger work lists run persistent, each wave of a capped
// grid striding over the work items (the kernels' persistent mode).  The
// environment variable TLNO_MAX_GRID_WAVES lowers the cap (read at every
// launch; used by the tests to force the persistent path).
long long max_grid_waves() {
    long long cap = 1LL << 24;
    if (const char* v = std::getenv("TLNO_MAX_GRID_WAVES")) {
        const long long n = std::atoll(v);
        if (n > 0) cap = std::min(cap, n);
    }
    return cap;
}

// nwaves waves per chunk, every chunk group of the class (FP32 or FP64
// kernels), or the team kernel with one 64-thread block per work item
void launch_jk(const ClassKernels& k, bool fp32, JKArgs& args, long long nwaves, hipStream_t st,
               int team_min_out = 0) {
    void* params[] = {&args};
    const long long cap = max_grid_waves();
    if (use_team(k, fp32, args, team_min_out)) {  // one block per shell quartet and slice (64 per task)
        const long long ns = std::max(1, args.team_slices);
        const long long items = args.tquart ? args.ntquart : static_cast<long long>(args.ntasks) * 64;
        if (items <= 0) return;
        const long long work = items * ns;
        args.stride = work > cap ? static_cast<int>(cap) : 0;
        HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(k.team),
                               dim3(static_cast<unsigned>(std::min(work, cap))), dim3(64), params, 0, st));
        args.stride = 0;
        return;
    }
    if (nwaves <= 0) return;
    if (nwaves > 0x7fffffffLL) throw std::runtime_error("tlno_gpu: class work list too large (more than 2^31 waves)");
    const int saved_stride = args.stride, saved_nwork = args.nwork;
    long long gw = nwaves;
    if (args.stride == 0 && nwaves > cap) {  // persistent: capped grid striding over the work items
        args.stride = static_cast<int>(cap);
        args.nwork = static_cast<int>(nwaves);
        gw = cap;
    }
    const unsigned grid = grid_for(gw * 64, 256);
    for (int g = 0; g < k.ngroup; ++g) {
        const bool fs = args.fs_min > 0.0;
        const JKKernel kern = fp32 ? (fs ? k.grp[g].jk32fs : k.grp[g].jk32) : (fs ? k.grp[g].jkfs : k.grp[g].jk);
        HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(kern), dim3(grid, k.grp[g].n), dim3(256), params, 0,
                               st));
    }
    args.stride = saved_stride;
    args.nwork = saved_nwork;
}

// Scratch for the g classes whose HRR and digestion work in global memory
// (HighKernels::gwork doubles per workgroup): one slot per CU, allocated on
// first use and kept with the plan.
void ensure_gwork(tlno_gpu_plan* pl, long long per_slot) {
    if (per_slot <= 0) return;
    if (pl->gslots == 0) {
        int ncu = 0;
        HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, pl->device));
        pl->gslots = std::max(1, ncu);
    }
    const long long need = per_slot * pl->gslots;
    if (need <= pl->gwork_doubles) return;
    require_alloc("global scratch of the g classes");
    if (pl->d_gwork) HIP_OK(hipFree(pl->d_gwork));
    pl->d_gwork = nullptr;
    pl->gwork_doubles = 0;
    HIP_OK(hipMalloc(&pl->d_gwork, static_cast<std::size_t>(need) * sizeof(double)));
    pl->gwork_doubles = need;
}

// team kernel of a class pair with an f or g shell over its quartet list
// (one workgroup of k.threads per quartet and slice; capped grids stride).
// Classes whose HRR and digestion work in global memory take one scratch
// slot per workgroup: the grid is capped at gslots workgroups striding over
// the quartets, and the caller runs such launches one after another (they
// share the scratch).
void launch_high(const HighKernels& k, JKArgs& args, hipStream_t st, double* gwork, int gslots) {
    void* params[] = {&args};
    const long long ns = std::max(1, args.team_slices);
    const long long work = args.ntquart * ns;
    if (work <= 0) return;
    long long cap = max_grid_waves();
    if (k.gwork > 0) {
        if (!gwork || gslots <= 0) throw std::runtime_error("tlno_gpu: g-class scratch not prepared");
        cap = std::min<long long>(cap, gslots);
        args.gwork = gwork;
    }
    args.stride = work > cap ? static_cast<int>(cap) : 0;
    HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(k.team), dim3(static_cast<unsigned>(std::min(work, cap))),
                           dim3(k.threads), params, 0, st));
    args.stride = 0;
}

// ---- density screen (per build) ----
// gmax[g] = max |D[i][j]| over the rows i of shell g (caller's basis)
__global__ void k_group_rowmax(const double* __restrict__ D, const int* row_group, int M,
                               unsigned long long* gmax) {
    const int i = blockIdx.x;
    double m = 0.0;
    for (int j = threadIdx.x; j < M; j += blockDim.x) m = fmax(m, fabs(D[(long)i * M + j]));
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = fmax(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0)
        atomicMax(gmax + row_group[i], static_cast<unsigned long long>(__double_as_longlong(m)));
}

// pairfac[P] = max(gmax[A], gmax[B]); clsmax[class] = max over the class
__global__ void k_pair_factor(const unsigned long long* gmax, const int* gA, const int* gB, const int* pcls,
                              int npair, double* pairfac, unsigned long long* clsmax) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = P < npair;
    const double g = in ? fmax(__longlong_as_double(gmax[gA[P]]), __longlong_as_double(gmax[gB[P]])) : 0.0;
    const int c = in ? pcls[P] : -1;
    if (in) pairfac[P] = g;
    // pairs are in class order: a wave whose pairs share one class takes
    // their maximum first (one atomic per wave, not per pair, on six slots)
    const int c0 = __shfl(c, 0, 64);
    if (__all(c == c0 || c < 0)) {
        double m = g;
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) m = fmax(m, __shfl_xor(m, sh, 64));
        if ((threadIdx.x & 63) == 0 && c0 >= 0)
            atomicMax(clsmax + c0, static_cast<unsigned long long>(__double_as_longlong(m)));
    } else if (in) {
        atomicMax(clsmax + c, static_cast<unsigned long long>(__double_as_longlong(g)));
    }
}

// cut[6 P + Y]: end of the ket prefix of class Y kept for bra pair P, from
// Q_P Q_q 4 max(g(P), max_Y g) >= tau — an upper bound of the reference's
// CVHF factor max(4 max(rho_ab, rho_cd), cross blocks), since every density
// block touching the quartet lies in the rows of one of its four shells.
__global__ void k_density_cut(const double* Q, const int* pcls, const int* cls_begin, const double* pairfac,
                              const unsigned long long* clsmax, int npair, double tau, int* cut) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= npair) return;
    const int X = pcls[P];
    for (int Y = 0; Y <= X; ++Y) {
        const double f = 4.0 * fmax(pairfac[P], __longlong_as_double(clsmax[Y]));
        // (1 + 1e-12): rounding never moves the prefix below a quartet the
        // exact per-quartet test keeps
        const double qp = Q[P] * f * (1.0 + 1e-12);
        int lo = cls_begin[Y], hi = cls_begin[Y + 1];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (qp * Q[mid] >= tau) lo = mid + 1;
            else hi = mid;
        }
        cut[6 * P + Y] = lo;
    }
}

// max |x| over n entries into *out (zeroed by the caller); non-negative
// doubles order like their bit patterns, so an integer max suffices
__global__ void k_absmax(const double* __restrict__ x, long n, unsigned long long* out) {
    double m = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = fmax(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// blk[gi][gj] = max |D| over the rows of shell gi and columns of shell gj and
// the transposed block (the reference's group_pair_density_max,
// integrals/two_electron/fock/detail/direct_build.ipp:126-178);
// bsum[gi][gj] = the sum over the block of the larger of |D_ab|, |D_ba|
// (bounds what one shell quartet adds to a J or K element per unit of its
// largest integral; used by the learned screen)
__global__ void k_block_max(const double* __restrict__ D, const int* __restrict__ grows, int ng, int M,
                            double* __restrict__ blk, double* __restrict__ bsum) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(ng) * ng) return;
    const int gi = static_cast<int>(t / ng), gj = static_cast<int>(t - static_cast<long long>(gi) * ng);
    double m = 0.0, sum = 0.0;
    for (int a = 0; a < kMaxComp; ++a) {
        const int ra = grows[kMaxComp * gi + a];
        if (ra < 0) break;
        for (int b = 0; b < kMaxComp; ++b) {
            const int rb = grows[kMaxComp * gj + b];
            if (rb < 0) break;
            const double v = fmax(fabs(D[static_cast<long long>(ra) * M + rb]), fabs(D[static_cast<long long>(rb) * M + ra]));
            m = fmax(m, v);
            sum += v;
        }
    }
    blk[t] = m;
    bsum[t] = sum;
}

// ---- per-quartet routing ----
// One wave per row {P, k_begin, k_end, 6 X + Y} of candidate kets (Schwarz
// survivors, cut further by the density prefix bound when that is on).  Each
// candidate is dropped, sent to the FP32 route or to the FP64 route:
//   * density test (the reference's rys_density_shell_quartet_live,
//     integrals/two_electron/rys/rys_engine.ipp:1124-1167): keep iff
//     est * max(4 max(rho_ab, rho_cd), max(D_ca, D_da, D_cb, D_db)) >= tau
//     (J only: the Coulomb term, K only: the exchange term);
//   * learned screen: est = Q_ab Q_cd min(1, 10^(y + margin)) with y the
//     network's prediction of log10(max|(ab|cd)| / (Q_ab Q_cd)); est must
//     reach the Schwarz threshold and, with the density test on, pass it
//     again with block sums of |D| in place of block maxima (est estimates
//     one integral, a block of them adds up); without it est = Q_ab Q_cd;
//   * precision: FP32 iff est * max|D| < gate.
// Survivors are compacted into the row's qlist region (FP64 from the front,
// FP32 from the back) and emitted as list-mode tasks of up to 64 kets into
// the row's own task slots (ceil(row length / 64) in t64 and in t32, fixed
// at plan time; unused slots get empty tasks), so no task counter is shared
// between rows and the host launches the class kernels without reading
// anything back.
constexpr int kRouteSeg = 512;
struct RouteArgs {
    const int4* rows;
    const long long* rbase;
    long long nrows;
    const int* cut;     // [6 npair] density prefix bound (nullptr: off)
    const double* Q;
    const int* pgA;
    const int* pgB;
    int ng;
    const double* blk;  // shell-block max |D| (nullptr: no density test)
    const double* bsum; // shell-block sum |D|
    int mode;           // 0 J and K, 1 J only, 2 K only
    double tau;
    double thr;
    MlpPacked mlp;
    FeatArgs fa;
    double margin;
    double gate;
    const double* maxD;
    int* qlist;
    int4* t64;
    int4* t32;
    const long long* rtask;       // first task slot of each row (in t64 and in t32)
    const double* decay;          // per pair decay indicator (nullptr: no decay test)
    double decay_thr;
    unsigned long long fp32_cls;  // bit 6 X + Y: class pair whose FP32 share runs the FP32 kernels
    int* cand;                    // learned builds: density survivors of each row (at its base)
    int* rcnt;                    // and their number per row
    unsigned long long* stats;    // [kCounterSlots][4]: candidates, density-kept, kept, FP32
};

__device__ inline int4 list_task(int P, long long off, int cnt) {
    return make_int4(P, static_cast<int>(static_cast<unsigned>(off & 0xffffffffll)),
                     cnt | static_cast<int>((off >> 32) << 8), 0);
}

// A row's list tasks in its fixed slots: n64 FP64 kets from the front of its
// qlist region, n32 FP32 kets from the back, 64 per task; unused slots get
// empty tasks (they leave at once).
__device__ inline void route_tasks(const RouteArgs& a, long long row, int P, long long base, long long len,
                                   int n64, int n32, bool cls32) {
    const int lane = threadIdx.x & 63;
    const long long slot = a.rtask[row];
    const int ncap = static_cast<int>((len + 63) >> 6);
    for (int i = lane; i < ncap; i += 64) {
        const int c64 = max(0, min(64, n64 - 64 * i));
        a.t64[slot + i] = list_task(P, base + 64ll * i, c64);
        if (cls32) {
            const int c32 = max(0, min(64, n32 - 64 * i));
            a.t32[slot + i] = list_task(P, base + len - 64ll * i - c32, c32);
        }
    }
}

// Per-row density (and decay) test.  One wave per row of candidate kets.
// CAND = false: the final routing of builds without the learned screen --
// kept kets compacted into the row's qlist region (FP64 from the front, FP32
// from the back), list tasks into the row's fixed task slots.
// CAND = true: the first pass of a learned build -- kept kets compacted into
// the row's candidate region (a.cand at the row's base), their number into
// a.rcnt[row]; k_route_net runs the network on them.  This pass holds no
// network registers, so it runs at full occupancy, which its dependent
// gathers (Q, pair groups, shell-block maxima) need.
template <bool CAND>
__global__ __launch_bounds__(256) void k_route(RouteArgs a) {
    const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= a.nrows) return;  // whole wave leaves together
    const int4 r = a.rows[wave];
    const int P = __builtin_amdgcn_readfirstlane(r.x);
    const int k0 = __builtin_amdgcn_readfirstlane(r.y);
    const int kend = __builtin_amdgcn_readfirstlane(r.z);
    const int xy = __builtin_amdgcn_readfirstlane(r.w);
    const int Y = xy - 6 * (xy / 6);
    int qe = kend;
    if (a.cut) qe = min(qe, a.cut[6 * P + Y]);
    const long long base = a.rbase[wave];
    const long long len = kend - k0;
    const double QP = a.Q[P];
    const int ga = a.pgA[P], gb = a.pgB[P];
    const long long ng = a.ng;
    const double rhoP = a.blk ? a.blk[ga * ng + gb] : 0.0;
    const double maxD = a.gate > 0.0 ? *a.maxD : 0.0;
    // classes whose FP32 share would run FP64 kernels anyway (team evaluator,
    // or FP32 kernels measured slower) keep every quartet in the FP64 list:
    // one list fills its tasks better than two
    const bool cls32 = !CAND && a.gate > 0.0 && ((a.fp32_cls >> xy) & 1ull);
    const unsigned long long below = (1ull << lane) - 1ull;
    int n64 = 0, n32 = 0;
    unsigned long long cand = 0, dkept = 0;
    for (int q0 = k0; q0 < qe; q0 += 64) {
        const int Qr = q0 + lane;
        const bool valid = Qr < qe;
        const int Qi = valid ? Qr : qe - 1;
        const double qq = QP * a.Q[Qi];
        bool keep = valid;
        if (a.blk) {
            const long long gc = a.pgA[Qi], gd = a.pgB[Qi];
            const double coul = 4.0 * fmax(rhoP, a.blk[gc * ng + gd]);
            const double exch = fmax(fmax(a.blk[gc * ng + ga], a.blk[gd * ng + ga]),
                                     fmax(a.blk[gc * ng + gb], a.blk[gd * ng + gb]));
            const double fac = a.mode == 1 ? coul : (a.mode == 2 ? exch : fmax(coul, exch));
            keep = keep && qq * fac >= a.tau;
        }
        if (a.decay) keep = keep && a.decay[P] * a.decay[Qi] >= a.decay_thr;
        cand += valid ? 1 : 0;
        dkept += keep ? 1 : 0;
        if constexpr (CAND) {
            const unsigned long long m = __ballot(keep);
            if (keep) a.cand[base + n64 + __popcll(m & below)] = Qr;
            n64 += __popcll(m);
        } else {
            const bool to32 = keep && cls32 && qq * maxD < a.gate;
            const bool to64 = keep && !to32;
            const unsigned long long m64 = __ballot(to64), m32 = __ballot(to32);
            if (to64) a.qlist[base + n64 + __popcll(m64 & below)] = Qr;
            if (to32) a.qlist[base + len - 1 - (n32 + __popcll(m32 & below))] = Qr;
            n64 += __popcll(m64);
            n32 += __popcll(m32);
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        cand += __shfl_xor(cand, m, 64);
        dkept += __shfl_xor(dkept, m, 64);
    }
    unsigned long long* st = a.stats + 4 * (wave % kCounterSlots);
    if constexpr (CAND) {
        if (lane == 0) {
            a.rcnt[wave] = n64;
            atomicAdd(st + 0, cand);
            atomicAdd(st + 1, dkept);
        }
    } else {
        route_tasks(a, wave, P, base, len, n64, n32, cls32);
        if (lane == 0) {
            atomicAdd(st + 0, cand);
            atomicAdd(st + 1, dkept);
            atomicAdd(st + 2, static_cast<unsigned long long>(n64 + n32));
            atomicAdd(st + 3, static_cast<unsigned long long>(n32));
        }
    }
}

// Second pass of a learned build: one wave per row, over the row's density
// survivors (a.cand) in dense chunks of 64.  Each chunk's 32 features go
// through the network on the matrix cores; est = Q_ab Q_cd min(1,
// 10^(y + margin)) must reach the Schwarz threshold and pass the density
// test again with shell-block sums of |D| in place of maxima (est estimates
// one integral; a block adds up to the sum); FP32 iff est max|D| < gate.
// Survivors compacted into the row's qlist region, list tasks into its slots.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_route_net(RouteArgs a) {
    // the network's packed weights, read by every chunk, sit in LDS (13 KB per
    // block) instead of being fetched from the cache hierarchy per chunk
    __shared__ uint4 sfrag[MLP_NFRAG * 64];
    __shared__ float svec[192];
    for (int i = threadIdx.x; i < MLP_NFRAG * 64; i += blockDim.x) sfrag[i] = a.mlp.frag[i];
    for (int i = threadIdx.x; i < 192; i += blockDim.x) svec[i] = a.mlp.vec[i];
    __syncthreads();
    const MlpPacked mlp{sfrag, svec, a.mlp.norm, a.mlp.b3};
    const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= a.nrows) return;  // whole wave leaves together
    const int4 r = a.rows[wave];
    const int P = __builtin_amdgcn_readfirstlane(r.x);
    const int xy = __builtin_amdgcn_readfirstlane(r.w);
    const int X = xy / 6, Y = xy - 6 * (xy / 6);
    const long long base = a.rbase[wave];
    const long long len = r.z - r.y;
    const int n = __builtin_amdgcn_readfirstlane(a.rcnt[wave]);
    const double QP = a.Q[P];
    const int ga = a.pgA[P], gb = a.pgB[P];
    const long long ng = a.ng;
    const double sigP = a.blk ? a.bsum[ga * ng + gb] : 0.0;
    const double maxD = a.gate > 0.0 ? *a.maxD : 0.0;
    const bool cls32 = a.gate > 0.0 && ((a.fp32_cls >> xy) & 1ull);
    const unsigned long long below = (1ull << lane) - 1ull;
    int n64 = 0, n32 = 0;
    for (int c0 = 0; c0 < n; c0 += 64) {
        const bool valid = c0 + lane < n;
        const int Qr = a.cand[base + (valid ? c0 + lane : c0)];  // invalid lanes run a stand-in quartet
        const double qq = QP * a.Q[Qr];
        float x[MLP_IN];
        quartet_features(a.fa, P, Qr, X, Y, x);
        const float y = mlp_forward_wave(mlp, x);  // every lane of the wave
        // 10^(y + margin) in single precision (the screen's estimate, not a
        // result: single-precision rounding of the factor is immaterial)
        const double est = qq * fminf(1.0f, exp2f(3.3219280948873623f * (y + static_cast<float>(a.margin))));
        bool keep = valid && est >= a.thr;
        if (a.blk) {
            const long long gc = a.pgA[Qr], gd = a.pgB[Qr];
            const double coul = 4.0 * fmax(sigP, a.bsum[gc * ng + gd]);
            const double exch = fmax(fmax(a.bsum[gc * ng + ga], a.bsum[gd * ng + ga]),
                                     fmax(a.bsum[gc * ng + gb], a.bsum[gd * ng + gb]));
            const double fac = a.mode == 1 ? coul : (a.mode == 2 ? exch : fmax(coul, exch));
            keep = keep && est * fac >= a.tau;
        }
        const bool to32 = keep && cls32 && est * maxD < a.gate;
        const bool to64 = keep && !to32;
        const unsigned long long m64 = __ballot(to64), m32 = __ballot(to32);
        if (to64) a.qlist[base + n64 + __popcll(m64 & below)] = Qr;
        if (to32) a.qlist[base + len - 1 - (n32 + __popcll(m32 & below))] = Qr;
        n64 += __popcll(m64);
        n32 += __popcll(m32);
    }
    route_tasks(a, wave, P, base, len, n64, n32, cls32);
    if (lane == 0) {
        unsigned long long* st = a.stats + 4 * (wave % kCounterSlots);
        atomicAdd(st + 2, static_cast<unsigned long long>(n64 + n32));
        atomicAdd(st + 3, static_cast<unsigned long long>(n32));
    }
}

// features of n quartets (pair_p[i] | pair_q[i]); out[32 i ..]
__global__ void k_features(FeatArgs fa, const int* __restrict__ pcls, const int* __restrict__ pp,
                           const int* __restrict__ pq, int n, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[MLP_IN];
    quartet_features(fa, pp[i], pq[i], pcls[pp[i]], pcls[pq[i]], x);
    for (int k = 0; k < MLP_IN; ++k) out[static_cast<long long>(i) * MLP_IN + k] = x[k];
}

void launch_store(const ClassKernels& k, StoreArgs& args, hipStream_t st) {
    void* params[] = {&args};
    HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(k.store), dim3(grid_for(args.n, 128), k.nchunk),
                           dim3(128), params, 0, st));
}

// Evaluates full Cartesian pair-orientation blocks for a list of (P, Q)
// quartets that all belong to class pair (X, Y), X >= Y.
void run_store(tlno_gpu_plan* pl, int X, int Y, const std::vector<int2>& qs, const std::vector<long long>& off,
               double* d_out) {
    if (qs.empty()) return;
    DeviceCtx& ctx = device_ctx(pl->device);
    int2* dq = nullptr;
    long long* doff = nullptr;
    HIP_OK(hipMalloc(&dq, qs.size() * sizeof(int2)));
    HIP_OK(hipMalloc(&doff, off.size() * sizeof(long long)));
    HIP_OK(hipMemcpyAsync(dq, qs.data(), qs.size() * sizeof(int2), hipMemcpyHostToDevice, pl->stream));
    HIP_OK(hipMemcpyAsync(doff, off.data(), off.size() * sizeof(long long), hipMemcpyHostToDevice, pl->stream));
    StoreArgs a{};
    a.pd = pl->pair_data();
    a.tb = ctx.tb;
    a.quartets = dq;
    a.off = doff;
    a.n = static_cast<int>(qs.size());
    a.out = d_out;
    if (X >= kNumLowClasses) {
        // classes with an f or g shell: team-evaluator blocks, one workgroup
        // per quartet (a scratch slot each where the working set is global)
        void* params[] = {&a};
        const HighKernels& hk = g_high[X - kNumLowClasses][Y];
        long long nb = std::min<long long>(a.n, max_grid_waves());
        if (hk.gwork > 0) {
            ensure_gwork(pl, hk.gwork);
            a.gwork = pl->d_gwork;
            nb = std::min<long long>(nb, pl->gslots);
        }
        HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(hk.store), dim3(static_cast<unsigned>(nb)),
                               dim3(hk.threads), params, 0, pl->stream));
    } else {
        launch_store(g_kernels[X][Y], a, pl->stream);
    }
    HIP_OK(hipStreamSynchronize(pl->stream));
    HIP_OK(hipFree(dq));
    HIP_OK(hipFree(doff));
}

void upload_pairs(tlno_gpu_plan* pl, const std::vector<int>& order,
                  const std::vector<std::vector<double4>>& prims_P, const std::vector<std::vector<double>>& prims_w,
                  const std::vector<std::vector<double>>& prims_ab) {
    const int np = static_cast<int>(order.size());
    std::vector<double> A(3 * np), AB(3 * np), pw, pab;
    std::vector<int> ao(2 * np), poff(np), pn(np), diag(np), pL(2 * np);
    std::vector<double4> pP;
    std::vector<int> gA(np), gB(np);
    for (int k = 0; k < np; ++k) {
        const int src = order[k];
        const GroupH& Ga = pl->groups[pl->pgA[src]];
        const GroupH& Gb = pl->groups[pl->pgB[src]];
        for (int x = 0; x < 3; ++x) {
            A[3 * k + x] = Ga.C[x];
            AB[3 * k + x] = Ga.C[x] - Gb.C[x];
        }
        ao[2 * k] = Ga.int_base;
        ao[2 * k + 1] = Gb.int_base;
        poff[k] = static_cast<int>(pP.size());
        pn[k] = static_cast<int>(prims_P[src].size());
        diag[k] = pl->pgA[src] == pl->pgB[src];
        pP.insert(pP.end(), prims_P[src].begin(), prims_P[src].end());
        pw.insert(pw.end(), prims_w[src].begin(), prims_w[src].end());
        pab.insert(pab.end(), prims_ab[src].begin(), prims_ab[src].end());
        pL[2 * k] = Ga.L;
        pL[2 * k + 1] = Gb.L;
        gA[k] = pl->pgA[src];
        gB[k] = pl->pgB[src];
    }
    pl->pgA = gA;
    pl->pgB = gB;
    pl->pn_host = pn;
    pl->nprim_pairs = static_cast<long long>(pP.size());
    // host mirror (molecule batches concatenate the pair data of their plans)
    pl->h_A = A;
    pl->h_AB = AB;
    pl->h_ao = ao;
    pl->h_poff = poff;
    pl->h_diag = diag;
    pl->h_pP = pP;
    pl->h_pw = pw;
    dev_upload(pl->dA, A);
    dev_upload(pl->dAB, AB);
    dev_upload(pl->dao, ao);
    dev_upload(pl->dpoff, poff);
    dev_upload(pl->dpn, pn);
    dev_upload(pl->ddiag, diag);
    dev_upload(pl->dpP, pP);
    dev_upload(pl->dpw, pw);
    dev_upload(pl->dpexp, pab);
    dev_upload(pl->dpL, pL);
}

// Features of each shell pair for the learned screen (plan order):
//   f0 = {log10 Q, log10 p_min, log10 p_max, log10 mu}, mu and P taken from the
//   primitive pair with the smallest p (the most diffuse one);
//   f1 = {|AB|, mu |AB|^2, log10 sum_prim |K c_a c_b / p|, log10 n_prim_pairs};
//   cen = {P, p_min}.
void upload_pair_features(tlno_gpu_plan* pl, const std::vector<int>& order,
                          const std::vector<std::vector<double4>>& prims_P,
                          const std::vector<std::vector<double>>& prims_w,
                          const std::vector<std::vector<double>>& prims_ab) {
    const int np = static_cast<int>(order.size());
    std::vector<float4> f0(np), f1(np), cen(np);
    auto lg = [](double v) { return static_cast<float>(std::log10(std::max(v, 1e-300))); };
    for (int k = 0; k < np; ++k) {
        const int src = order[k];
        const auto& P = prims_P[src];
        const GroupH& Ga = pl->groups[pl->pgA[k]];
        const GroupH& Gb = pl->groups[pl->pgB[k]];
        double r2 = 0.0;
        for (int x = 0; x < 3; ++x) r2 += (Ga.C[x] - Gb.C[x]) * (Ga.C[x] - Gb.C[x]);
        if (P.empty()) {
            f0[k] = make_float4(lg(pl->Q[k]), 0.0f, 0.0f, 0.0f);
            f1[k] = make_float4(static_cast<float>(std::sqrt(r2)), 0.0f, -300.0f, 0.0f);
            cen[k] = make_float4(static_cast<float>(Ga.C[0]), static_cast<float>(Ga.C[1]),
                                 static_cast<float>(Ga.C[2]), 1.0f);
            continue;
        }
        std::size_t imin = 0;
        double pmax = 0.0, wsum = 0.0;
        for (std::size_t i = 0; i < P.size(); ++i) {
            if (P[i].w < P[imin].w) imin = i;
            pmax = std::max(pmax, P[i].w);
            wsum += std::abs(prims_w[src][i]);
        }
        const double al = prims_ab[src][2 * imin], be = prims_ab[src][2 * imin + 1];
        const double pmin = P[imin].w, mu = al * be / pmin;
        f0[k] = make_float4(lg(pl->Q[k]), lg(pmin), lg(pmax), lg(mu));
        f1[k] = make_float4(static_cast<float>(std::sqrt(r2)), static_cast<float>(mu * r2), lg(wsum),
                            lg(static_cast<double>(P.size())));
        cen[k] = make_float4(static_cast<float>(P[imin].x), static_cast<float>(P[imin].y),
                             static_cast<float>(P[imin].z), static_cast<float>(pmin));
    }
    dev_upload(pl->d_pf0, f0);
    dev_upload(pl->d_pf1, f1);
    dev_upload(pl->d_pcen, cen);
}

void build_plan(tlno_gpu_plan* pl, int n, const double* centers, const int* ang, const int* nprim,
                const double* exps, const double* coefs) {
    std::memset(pl->team_sel, -1, sizeof(pl->team_sel));
    // input validation first: invalid bases fail the same way with or without a device
    pl->groups = build_groups(n, centers, ang, nprim, exps, coefs);
    DeviceCtx& ctx = device_ctx(pl->device);
    (void)ctx;
    HIP_OK(hipSetDevice(pl->device));
    HIP_OK(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
    const int ng = static_cast<int>(pl->groups.size());
    int ib = 0, sb = 0;
    for (auto& G : pl->groups) {
        G.int_base = ib;
        G.sph_base = sb;
        ib += ncart(G.L);
        sb += 2 * G.L + 1;
    }
    pl->Mcart = ib;
    pl->Msph = sb;
    if (n != ib) throw std::runtime_error("eri_ao_tensor_hgp: inconsistent component count.");
    if (sb > 65535 || ib > 65535)
        throw std::runtime_error("eri_ao_tensor_hgp: direct J/K pair tables assume M <= 65535.");
    pl->Mout = pl->spherical ? pl->Msph : pl->Mcart;
    pl->int2basis.resize(ib);
    std::vector<int> iL(ib), iloc(ib), isb(ib);
    for (const auto& G : pl->groups)
        for (int k = 0; k < ncart(G.L); ++k) {
            pl->int2basis[G.int_base + k] = G.comp[k];
            iL[G.int_base + k] = G.L;
            iloc[G.int_base + k] = k;
            isb[G.int_base + k] = G.sph_base;
        }
    std::vector<int> sL(sb), sloc(sb), sib(sb);
    for (const auto& G : pl->groups)
        for (int s = 0; s < 2 * G.L + 1; ++s) {
            sL[G.sph_base + s] = G.L;
            sloc[G.sph_base + s] = s;
            sib[G.sph_base + s] = G.int_base;
        }
    dev_upload(pl->d_int2basis, pl->int2basis);
    dev_upload(pl->d_intL, iL);
    dev_upload(pl->d_intloc, iloc);
    dev_upload(pl->d_intsb, isb);
    dev_upload(pl->d_sphL, sL);
    dev_upload(pl->d_sphloc, sloc);
    dev_upload(pl->d_sphib, sib);

    // shell pairs, oriented with the higher L first (rys_engine.ipp:2443-2460)
    std::vector<int> pcls;
    std::vector<std::vector<double4>> prims_P;
    std::vector<std::vector<double>> prims_w, prims_ab;
    for (int lg = 0; lg < ng; ++lg)
        for (int rg = 0; rg <= lg; ++rg) {
            int a = lg, b = rg;
            if (pl->groups[b].L > pl->groups[a].L) std::swap(a, b);
            const GroupH& Ga = pl->groups[a];
            const GroupH& Gb = pl->groups[b];
            pl->pgA.push_back(a);
            pl->pgB.push_back(b);
            pcls.push_back(class_of(Ga.L, Gb.L));
            double AB[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                AB[x] = Ga.C[x] - Gb.C[x];
                r2 += AB[x] * AB[x];
            }
            std::vector<double4> P;
            std::vector<double> w, ab;
            for (int i = 0; i < Ga.nprim; ++i)
                for (int j = 0; j < Gb.nprim; ++j) {
                    const double al = Ga.exps[i], be = Gb.exps[j];
                    const double cab = Ga.cr[i] * Gb.cr[j];
                    if (cab == 0.0) continue;
                    const double p = al + be, ip = 1.0 / p;
                    const double K = std::exp(-(al * be * ip) * r2);
                    // P = (a A + b B)/p written as A - (b/p)(A - B): exact for
                    // one-centre pairs, so parity-forbidden integrals vanish.
                    double4 v;
                    v.x = Ga.C[0] - be * ip * AB[0];
                    v.y = Ga.C[1] - be * ip * AB[1];
                    v.z = Ga.C[2] - be * ip * AB[2];
                    v.w = p;
                    P.push_back(v);
                    w.push_back(K * cab * ip);
                    ab.push_back(al);
                    ab.push_back(be);
                }
            prims_P.push_back(std::move(P));
            prims_w.push_back(std::move(w));
            prims_ab.push_back(std::move(ab));
        }
    const int np = static_cast<int>(pcls.size());
    pl->npair = np;
    // order by class (stable)
    std::vector<int> order(np);
    for (int i = 0; i < np; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pcls[x] < pcls[y]; });
    std::vector<int> cls_sorted(np);
    for (int k = 0; k < np; ++k) cls_sorted[k] = pcls[order[k]];
    const std::vector<int> pgA0 = pl->pgA, pgB0 = pl->pgB;
    upload_pairs(pl, order, prims_P, prims_w, prims_ab);
    for (int c = 0; c <= kNumClasses; ++c)
        pl->cls_begin[c] =
            static_cast<int>(std::lower_bound(cls_sorted.begin(), cls_sorted.end(), c) - cls_sorted.begin());

    // Schwarz bounds from the diagonal blocks (ab|ab), on the GPU
    std::vector<double> Q(np, 0.0);
    for (int X = 0; X < kNumClasses; ++X) {
        const int b0 = pl->cls_begin[X], b1 = pl->cls_begin[X + 1];
        if (b1 <= b0) continue;
        const int LA = kClassL[X][0], LB = kClassL[X][1];
        const int na = ncart(LA), nb = ncart(LB), nab = na * nb;
        std::vector<int2> qs;
        std::vector<long long> off;
        for (int P = b0; P < b1; ++P) {
            qs.push_back(make_int2(P, P));
            off.push_back(static_cast<long long>(P - b0) * nab * nab);
        }
        double* dout = nullptr;
        const std::size_t total = static_cast<std::size_t>(b1 - b0) * nab * nab;
        HIP_OK(hipMalloc(&dout, total * sizeof(double)));
        run_store(pl, X, X, qs, off, dout);
        std::vector<double> h(total);
        HIP_OK(hipMemcpy(h.data(), dout, total * sizeof(double), hipMemcpyDeviceToHost));
        HIP_OK(hipFree(dout));
        const C2S& cs = c2s_table();
        for (int P = b0; P < b1; ++P) {
            const double* blk = h.data() + static_cast<std::size_t>(P - b0) * nab * nab;
            double mx = 0.0;
            if (!pl->spherical) {
                for (int ab = 0; ab < nab; ++ab) mx = std::max(mx, std::abs(blk[ab * nab + ab]));
            } else {
                const int sa = 2 * LA + 1, sbn = 2 * LB + 1;
                for (int s = 0; s < sa; ++s)
                    for (int t = 0; t < sbn; ++t) {
                        double v = 0.0;
                        for (int a = 0; a < na; ++a)
                            for (int b = 0; b < nb; ++b) {
                                const double wab = cs.c[LA][s][a] * cs.c[LB][t][b];
                                if (wab == 0.0) continue;
                                for (int c = 0; c < na; ++c)
                                    for (int d = 0; d < nb; ++d) {
                                        const double wcd = cs.c[LA][s][c] * cs.c[LB][t][d];
                                        if (wcd == 0.0) continue;
                                        v += wab * wcd * blk[(a * nb + b) * nab + c * nb + d];
                                    }
                            }
                        mx = std::max(mx, std::abs(v));
                    }
            }
            Q[P] = std::sqrt(mx);
        }
    }
    // sort inside each class by descending bound
    std::vector<int> order2(np);
    for (int i = 0; i < np; ++i) order2[i] = i;
    for (int c = 0; c < kNumClasses; ++c)
        std::stable_sort(order2.begin() + pl->cls_begin[c], order2.begin() + pl->cls_begin[c + 1],
                         [&](int x, int y) { return Q[x] > Q[y]; });
    // compose with the first ordering
    std::vector<int> final_order(np);
    for (int k = 0; k < np; ++k) final_order[k] = order[order2[k]];
    pl->pgA = pgA0;
    pl->pgB = pgB0;
    upload_pairs(pl, final_order, prims_P, prims_w, prims_ab);
    pl->Q.resize(np);
    for (int k = 0; k < np; ++k) pl->Q[k] = Q[order2[k]];
    // heuristic shell-pair decay indicator (rys_engine.ipp:2205-2241):
    // D_ab = K_a K_b exp(-mu_min |AB|^2), K_g = sum_i |c_i| (pi/a_i)^{3/4}
    // over the radial coefficients, mu_min from the two smallest exponents
    {
        std::vector<double> Kg(ng), amin(ng), dec(np);
        for (int g = 0; g < ng; ++g) {
            const GroupH& G = pl->groups[g];
            double k = 0.0, a0 = G.exps[0];
            for (int i = 0; i < G.nprim; ++i) {
                a0 = std::min(a0, G.exps[i]);
                k += std::abs(G.cr[i]) * std::pow(3.14159265358979323846 / G.exps[i], 0.75);
            }
            Kg[g] = k;
            amin[g] = a0;
        }
        for (int k = 0; k < np; ++k) {
            const int a = pl->pgA[k], b = pl->pgB[k];
            double r2 = 0.0;
            for (int x = 0; x < 3; ++x) r2 += (pl->groups[a].C[x] - pl->groups[b].C[x]) * (pl->groups[a].C[x] - pl->groups[b].C[x]);
            dec[k] = Kg[a] * Kg[b] * std::exp(-(amin[a] * amin[b] / (amin[a] + amin[b])) * r2);
        }
        dev_upload(pl->d_decay, dec);
    }

    const std::size_t mc2 = static_cast<std::size_t>(pl->Mcart) * pl->Mcart;
    const std::size_t mo2 = static_cast<std::size_t>(pl->Mout) * pl->Mout;
    HIP_OK(hipMalloc(&pl->dDin, mo2 * sizeof(double)));
    HIP_OK(hipMalloc(&pl->dDint, mc2 * sizeof(double)));
    HIP_OK(hipMalloc(&pl->dJacc, mc2 * sizeof(double)));
    HIP_OK(hipMalloc(&pl->dKacc, mc2 * sizeof(double)));
    HIP_OK(hipMalloc(&pl->dJout, mo2 * sizeof(double)));
    HIP_OK(hipMalloc(&pl->dKout, mo2 * sizeof(double)));
    HIP_OK(hipMalloc(&pl->d_scalar, 8 * sizeof(double)));
    upload_pair_features(pl, final_order, prims_P, prims_w, prims_ab);
    // density-screen metadata
    {
        std::vector<int> row_group(pl->Mout), pcls(np), clsb(kNumClasses + 1);
        for (int g = 0; g < ng; ++g) {
            const GroupH& G = pl->groups[g];
            if (pl->spherical)
                for (int s = 0; s < 2 * G.L + 1; ++s) row_group[G.sph_base + s] = g;
            else
                for (int k = 0; k < ncart(G.L); ++k) row_group[G.comp[k]] = g;
        }
        for (int c = 0; c < kNumClasses; ++c)
            for (int P = pl->cls_begin[c]; P < pl->cls_begin[c + 1]; ++P) pcls[P] = c;
        for (int c = 0; c <= kNumClasses; ++c) clsb[c] = pl->cls_begin[c];
        dev_upload(pl->d_row_group, row_group);
        std::vector<int> grows(kMaxComp * static_cast<std::size_t>(std::max(1, ng)), -1);
        for (int g = 0; g < ng; ++g) {
            const GroupH& G = pl->groups[g];
            if (pl->spherical)
                for (int s = 0; s < 2 * G.L + 1; ++s) grows[kMaxComp * g + s] = G.sph_base + s;
            else
                for (int k = 0; k < ncart(G.L); ++k) grows[kMaxComp * g + k] = G.comp[k];
        }
        dev_upload(pl->d_grows, grows);
        HIP_OK(hipMalloc(&pl->d_blk, std::max<std::size_t>(1, static_cast<std::size_t>(ng) * ng) * sizeof(double)));
        HIP_OK(hipMalloc(&pl->d_bsum, std::max<std::size_t>(1, static_cast<std::size_t>(ng) * ng) * sizeof(double)));
        HIP_OK(hipMalloc(&pl->d_rstats, 4 * kCounterSlots * sizeof(unsigned long long)));
        dev_upload(pl->d_pgA, pl->pgA);
        dev_upload(pl->d_pgB, pl->pgB);
        dev_upload(pl->d_pcls, pcls);
        dev_upload(pl->d_clsb, clsb);
        dev_upload(pl->d_Q, pl->Q);
        HIP_OK(hipMalloc(&pl->d_gmax, std::max(1, ng) * sizeof(unsigned long long)));
        HIP_OK(hipMalloc(&pl->d_clsmax, 16 * sizeof(unsigned long long)));
        HIP_OK(hipMalloc(&pl->d_pairfac, std::max(1, np) * sizeof(double)));
        HIP_OK(hipMalloc(&pl->d_cut, 6 * static_cast<std::size_t>(std::max(1, np)) * sizeof(int)));
        HIP_OK(hipMalloc(&pl->d_counter, 4 * kCounterSlots * sizeof(unsigned long long)));
    }
}

// Waves per task for a class whose tasks should be split `split` ways: a task
// with nq <= 32 kets already spreads its bra primitive pairs over the 64 / ks
// lanes of each ket (ks = nq rounded up to a power of two; device side
// ket_stride), so it needs that many times fewer waves.
int wave_split(int split, int nq) {
    int ks = 1;
    while (ks < nq && ks < 64) ks <<= 1;
    const int lanes_per_ket = 64 / ks;
    return (split + lanes_per_ket - 1) / lanes_per_ket;
}

// Classes whose tasks give fewer waves than this split each task's bra
// primitive pairs over several waves.  Every slice digests its own partial
// block (one FP64 atomic per J/K element), so the split trades parallelism
// for memory-side atomics: 4096 measured best (organic30 18.2 ms at 16384,
// 15.3 ms at 4096, 16.1 ms without splitting).
#ifndef TLNO_TARGET_WAVES
#define TLNO_TARGET_WAVES 4096
#endif

// bucket of a primitive-pair count: half octaves, so counts inside one
// bucket differ by less than a factor sqrt(2)
inline int prim_bucket(int pn) {
    int b = 0;
    while ((1LL << (b + 1)) <= 2LL * pn * pn) ++b;  // b = floor(log2(2 pn^2)) = floor(2 log2 pn) + 1
    return b;
}

// Bucketed list tasks of class pair (X, Y) (see TaskSet).  bk: the ket
// groups of class Y, gbeg/gend their ranges in bk.
void bucket_list_tasks(const tlno_gpu_plan* pl, double thr, int X, int Y, const std::vector<int>& bk,
                       const std::vector<std::pair<int, int>>& groups, std::vector<int4>& out,
                       std::vector<double>& bound) {
    std::vector<std::pair<double, int4>> tasks;
    for (int P = pl->cls_begin[X]; P < pl->cls_begin[X + 1]; ++P)
        for (const auto& g : groups) {
            // group entries are pair indices in increasing order (bounds descending)
            int lo = g.first, hi = g.second;
            if (thr > 0.0) {
                int a = lo, b = hi;
                while (a < b) {
                    const int mid = (a + b) / 2;
                    if (pl->Q[P] * pl->Q[bk[mid]] >= thr) a = mid + 1;
                    else b = mid;
                }
                hi = a;
            }
            if (X == Y) {  // unique quartets: kets Q <= P
                int a = lo, b = hi;
                while (a < b) {
                    const int mid = (a + b) / 2;
                    if (bk[mid] <= P) a = mid + 1;
                    else b = mid;
                }
                hi = a;
            }
            for (int j0 = lo; j0 < hi; j0 += 64) {
                const long long off = j0;
                const int n = std::min(64, hi - j0);
                tasks.push_back({pl->Q[P] * pl->Q[bk[j0]],
                                 make_int4(P, static_cast<int>(off & 0xffffffffLL),
                                           n | static_cast<int>((off >> 32) << 8), 0)});
            }
        }
    // heaviest first; the tasks below a precision gate form a suffix
    std::stable_sort(tasks.begin(), tasks.end(),
                     [](const std::pair<double, int4>& u, const std::pair<double, int4>& v) { return u.first > v.first; });
    for (const auto& t : tasks) {
        out.push_back(t.second);
        bound.push_back(t.first);
    }
}

// Bra primitive slices of one class's list-mode tasks, balanced by cost.
// The lane evaluator launches fewer waves than the machine holds, so a
// launch lasts as long as its slowest wave, and a wave's time is that of its
// busiest lane: ceil(bra pairs of the slice / lanes per ket) x ket pairs (the
// flattened grid for flat_prims classes: ceil(bra x ket pairs / lanes per
// ket)).  The smallest per-wave cost C is chosen for which the slices
// (each task cut into the fewest equal slices costing at most C) number at
// most max(target, 5/4 of the tasks) waves over the class's chunks; each
// slice digests its own partial block, which bounds the split.  Slices stay
// adjacent, so the bound order of the list (the FP32 suffix) is kept.
// Measured on organic30 (autotuned build): 15.4 -> 14.65 ms; the classes'
// serialized launches 15-40 % shorter.
void balance_list_slices(const tlno_gpu_plan* pl, const std::vector<int>& bk, int X, int Y,
                         const std::vector<int4>& in, const std::vector<double>& in_bound, std::vector<int4>& out,
                         std::vector<double>& out_bound, std::vector<double>& out_cost) {
    const bool flat = flat_prims(kClassLAB[X], 0, kClassLAB[Y], 0);
    const long long nchunk = std::max(1, g_kernels[X][Y].nchunk);
    const std::size_t n = in.size();
    std::vector<int> pbn(n), pkn(n), ibs(n);
    for (std::size_t i = 0; i < n; ++i) {
        const int4& t = in[i];
        const long long off = static_cast<long long>(static_cast<unsigned>(t.y)) | (static_cast<long long>(t.z >> 8) << 32);
        const int nq = t.z & 255;
        int mk = 1;
        for (int j = 0; j < nq; ++j) mk = std::max(mk, pl->pn_host[bk[off + j]]);
        int ks = 1;
        while (ks < nq && ks < 64) ks <<= 1;
        pbn[i] = std::max(1, pl->pn_host[t.x]);
        pkn[i] = mk;
        ibs[i] = 64 / ks;
    }
    // fewest slices of task i whose busiest lane costs at most C
    auto slices = [&](std::size_t i, double C) {
        long long nbmax;  // bra pairs per slice within the cost
        if (flat) nbmax = static_cast<long long>(std::floor(C)) * ibs[i] / pkn[i];
        else nbmax = static_cast<long long>(std::floor(C / pkn[i])) * ibs[i];
        if (nbmax < 1) return pbn[i];
        return static_cast<int>(std::min<long long>(pbn[i], (pbn[i] + nbmax - 1) / nbmax));
    };
    auto cost1 = [&](std::size_t i) {
        return flat ? std::ceil(static_cast<double>(pbn[i]) * pkn[i] / ibs[i])
                    : std::ceil(static_cast<double>(pbn[i]) / ibs[i]) * pkn[i];
    };
    static const long long target = [] {
        const char* e = std::getenv("TLNO_LIST_TARGET_WAVES");
        return e ? std::max(1LL, std::atoll(e)) : static_cast<long long>(TLNO_TARGET_WAVES);
    }();
    // a launch of many times the target is balanced by the hardware's
    // dynamic dispatch: no split (water40 FP64 measured 0.6 % slower split)
    const long long nw = static_cast<long long>(n) * nchunk;
    const long long cap = nw >= 8 * target ? nw : std::max(target, nw * 5 / 4);
    double lo = 1.0, hi = 1.0;
    for (std::size_t i = 0; i < n; ++i) hi = std::max(hi, cost1(i));
    auto count = [&](double C) {
        long long w = 0;
        for (std::size_t i = 0; i < n; ++i) w += slices(i, C);
        return w * nchunk;
    };
    if (count(lo) > cap) {
        for (int it = 0; it < 48 && hi - lo > 0.5; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (count(mid) > cap) lo = mid;
            else hi = mid;
        }
    } else {
        hi = lo;
    }
    for (std::size_t i = 0; i < n; ++i) {
        const int s = slices(i, hi);
        const int step = (pbn[i] + s - 1) / s;
        for (int i0 = 0; i0 < pbn[i]; i0 += step) {
            const int i1 = std::min(pbn[i], i0 + step);
            int4 t = in[i];
            t.w = i0 | (i1 << 16);
            out.push_back(t);
            out_bound.push_back(in_bound[i]);
            const double nbs = i1 - i0;
            out_cost.push_back(flat ? std::ceil(nbs * pkn[i] / ibs[i]) : std::ceil(nbs / ibs[i]) * pkn[i]);
        }
    }
}

TaskSet& get_tasks(tlno_gpu_plan* pl, double thr) {
    auto it = pl->tasks.find(thr);
    if (it != pl->tasks.end()) return it->second;
    require_alloc("work lists for this Schwarz threshold");
    TaskSet ts;
    std::vector<int4> all;
    std::vector<long long> pref(static_cast<std::size_t>(pl->npair) + 1, 0);
    for (int k = 0; k < pl->npair; ++k) pref[k + 1] = pref[k] + pl->pn_host[k];
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            ts.off[X][Y] = static_cast<long long>(all.size());
            const int b0 = pl->cls_begin[X], b1 = pl->cls_begin[X + 1];
            const int k0 = pl->cls_begin[Y], k1 = pl->cls_begin[Y + 1];
            if (b1 > b0 && k1 > k0) {
                std::vector<int4> base;
                for (int P = b0; P < b1; ++P) {
                    int end = k1;
                    if (thr > 0.0) {
                        // largest prefix with Q_P * Q_Q >= thr (bounds sorted descending)
                        int lo = k0, hi = k1;
                        while (lo < hi) {
                            const int mid = (lo + hi) / 2;
                            if (pl->Q[P] * pl->Q[mid] >= thr) lo = mid + 1;
                            else hi = mid;
                        }
                        end = lo;
                    }
                    if (X == Y) end = std::min(end, P + 1);
                    for (int q0 = k0; q0 < end; q0 += 64) base.push_back(make_int4(P, q0, end, 0));
                    const long long nq = std::max(0, end - k0);
                    ts.shell_quartets += nq;
                    ts.cls_sq[X][Y] += nq;
                    if (end > k0) {
                        const long long pq = (pref[end] - pref[k0]) * pl->pn_host[P];
                        ts.prim_quartets += pq;
                        ts.cls_pq[X][Y] += pq;
                    }
                }
                // Too few waves to fill 256 CUs (rare, heavy classes): split the
                // bra primitive pairs of each task into slices of their own.
                const long long waves = static_cast<long long>(base.size()) * g_kernels[X][Y].nchunk;
                constexpr long long kTargetWaves = TLNO_TARGET_WAVES;
                int split = 1;
                if (waves > 0 && waves < kTargetWaves)
                    split = static_cast<int>(std::min<long long>(256, (kTargetWaves + waves - 1) / waves));
                std::vector<int4> cls;
                for (const int4& t : base) {
                    const int pbn = pl->pn_host[t.x];
                    const int s = std::max(1, std::min(pbn, wave_split(split, std::min(t.z - t.y, 64))));
                    const int step = (pbn + s - 1) / s;
                    for (int i0 = 0; i0 < pbn; i0 += step) {
                        const int i1 = std::min(pbn, i0 + step);
                        cls.push_back(make_int4(t.x, t.y, t.z, i0 | (i1 << 16)));
                    }
                }
                // Order the class's tasks by their largest bound Q_P Q_q_begin
                // (kets are bound-sorted, so q_begin carries the task's max):
                // heavy tasks start first, and the tasks below any gate on
                // bound x max|D| form a suffix (the FP32 share of the
                // mixed-precision route).
                std::stable_sort(cls.begin(), cls.end(), [&](const int4& u, const int4& v) {
                    return pl->Q[u.x] * pl->Q[u.y] > pl->Q[v.x] * pl->Q[v.y];
                });
                auto& bm = ts.bmax[X][Y];
                auto& csq = ts.cum_sq[X][Y];
                auto& cpq = ts.cum_pq[X][Y];
                csq.push_back(0);
                cpq.push_back(0);
                for (const int4& t : cls) {
                    bm.push_back(pl->Q[t.x] * pl->Q[t.y]);
                    const int qe = std::min(t.z, t.y + 64);
                    const long long ib0 = t.w & 0xffff, ib1 = (t.w >> 16) & 0xffff;
                    csq.push_back(csq.back() + (ib0 == 0 ? qe - t.y : 0));
                    cpq.push_back(cpq.back() + (pref[qe] - pref[t.y]) * (ib1 - ib0));
                    all.push_back(t);
                }
            }
            ts.cnt[X][Y] = static_cast<long long>(all.size()) - ts.off[X][Y];
        }
    if (!all.empty()) {
        HIP_OK(hipMalloc(&ts.d_tasks, all.size() * sizeof(int4)));
        HIP_OK(hipMemcpy(ts.d_tasks, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    // ket groups by primitive-pair bucket, for classes with more than one bucket
    std::vector<int> bk;
    std::vector<std::pair<int, int>> kgroups[6];
    for (int c = 0; c < 6; ++c) {
        std::map<int, std::vector<int>> by;
        for (int Q = pl->cls_begin[c]; Q < pl->cls_begin[c + 1]; ++Q) by[prim_bucket(pl->pn_host[Q])].push_back(Q);
        if (by.size() < 2) continue;
        for (auto& kv : by) {
            const int b = static_cast<int>(bk.size());
            bk.insert(bk.end(), kv.second.begin(), kv.second.end());
            kgroups[c].push_back({b, static_cast<int>(bk.size())});
        }
    }
    std::vector<int4> lall, lpt;
    std::vector<double> lcost;
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            ts.loff[X][Y] = static_cast<long long>(lall.size());
            if (!kgroups[Y].empty() && ts.cls_sq[X][Y] > 0) {
                std::vector<int4> cls;
                std::vector<double> bound;
                bucket_list_tasks(pl, thr, X, Y, bk, kgroups[Y], cls, bound);
                balance_list_slices(pl, bk, X, Y, cls, bound, lall, ts.lbmax[X][Y], lcost);
            }
            ts.lcnt[X][Y] = static_cast<long long>(lall.size()) - ts.loff[X][Y];
            // the same tasks heaviest first, for builds that run the class's
            // whole list on one kernel (no FP32 suffix to keep at the end)
            std::vector<std::size_t> ord(static_cast<std::size_t>(ts.lcnt[X][Y]));
            for (std::size_t i = 0; i < ord.size(); ++i) ord[i] = static_cast<std::size_t>(ts.loff[X][Y]) + i;
            std::stable_sort(ord.begin(), ord.end(), [&](std::size_t u, std::size_t v) { return lcost[u] > lcost[v]; });
            for (std::size_t i : ord) lpt.push_back(lall[i]);
        }
    if (!lall.empty()) {
        HIP_OK(hipMalloc(&ts.d_bk, bk.size() * sizeof(int)));
        HIP_OK(hipMemcpy(ts.d_bk, bk.data(), bk.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_OK(hipMalloc(&ts.d_ltasks, 2 * lall.size() * sizeof(int4)));
        HIP_OK(hipMemcpy(ts.d_ltasks, lall.data(), lall.size() * sizeof(int4), hipMemcpyHostToDevice));
        ts.d_ltasks_lpt = ts.d_ltasks + lall.size();
        HIP_OK(hipMemcpy(ts.d_ltasks_lpt, lpt.data(), lpt.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    // Quartets with an f shell: the bra is the pair of the higher class (the
    // later one in plan order when both are of one class), the kets surviving
    // Schwarz a prefix of the bound-ordered ket class; one list per class
    // pair, heaviest quartets first (the team evaluator's order)
    std::vector<int2> hall;
    for (int X = kNumLowClasses; X < kNumClasses; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const int h = X - kNumLowClasses;
            ts.hoff[h][Y] = static_cast<long long>(hall.size());
            const int k0 = pl->cls_begin[Y], k1 = pl->cls_begin[Y + 1];
            std::vector<int2> cls;
            for (int P = pl->cls_begin[X]; P < pl->cls_begin[X + 1] && k1 > k0; ++P) {
                int end = k1;
                if (thr > 0.0) {
                    int lo = k0, hi = k1;
                    while (lo < hi) {
                        const int mid = (lo + hi) / 2;
                        if (pl->Q[P] * pl->Q[mid] >= thr) lo = mid + 1;
                        else hi = mid;
                    }
                    end = lo;
                }
                if (X == Y) end = std::min(end, P + 1);
                for (int Q = k0; Q < end; ++Q) cls.push_back(make_int2(P, Q));
                if (end > k0) {
                    ts.shell_quartets += end - k0;
                    ts.prim_quartets += (pref[end] - pref[k0]) * pl->pn_host[P];
                    ts.hpq[h][Y] += (pref[end] - pref[k0]) * pl->pn_host[P];
                }
            }
            std::stable_sort(cls.begin(), cls.end(), [&](const int2& u, const int2& v) {
                return static_cast<long long>(pl->pn_host[u.x]) * pl->pn_host[u.y] >
                       static_cast<long long>(pl->pn_host[v.x]) * pl->pn_host[v.y];
            });
            hall.insert(hall.end(), cls.begin(), cls.end());
            ts.hcnt[h][Y] = static_cast<long long>(hall.size()) - ts.hoff[h][Y];
        }
    if (!hall.empty()) {
        HIP_OK(hipMalloc(&ts.d_hq, hall.size() * sizeof(int2)));
        HIP_OK(hipMemcpy(ts.d_hq, hall.data(), hall.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    return pl->tasks.emplace(thr, ts).first->second;
}

inline int class_team_min(const tlno_gpu_plan* pl, int X, int Y) {
    const int sel = pl->team_sel[X][Y];
    return sel < 0 ? pl->team_min_out : (sel ? 1 : 0);
}

// Team evaluator waves per shell quartet for a class with nq quartets: a class
// too small to give every SIMD a wave splits each quartet's primitive
// combinations over several waves (each repeats the transfer and digestion,
// atomics included, so the target stays low: 1024 measured best of 1024 to
// 4096).
#ifndef TLNO_TEAM_TARGET_WAVES
#define TLNO_TEAM_TARGET_WAVES 1024
#endif
#ifndef TLNO_TEAM_MAX_SLICES
#define TLNO_TEAM_MAX_SLICES 8
#endif
int team_slices_for(long long nq) {
    constexpr long long kTeamTargetWaves = TLNO_TEAM_TARGET_WAVES;
    if (nq <= 0 || nq >= kTeamTargetWaves) return 1;
    return static_cast<int>(std::min<long long>(TLNO_TEAM_MAX_SLICES, (kTeamTargetWaves + nq - 1) / nq));
}

// Candidate rows for the per-quartet routing kernel (once per threshold):
// every Schwarz-surviving (P | ket range) split into rows of <= kRouteSeg
// kets, each owning a qlist region of its length; per class pair the task
// capacity is the number of 64-ket chunks of its rows (either route).
void ensure_route(tlno_gpu_plan* pl, TaskSet& ts, double thr) {
    if (ts.d_rbase) return;
    require_alloc("routing buffers for this threshold");
    std::vector<int4> rows;
    std::vector<long long> rbase, rtask;
    long long q = 0, cap = 0;
    for (int X = 5; X >= 0; --X)
        for (int Y = X; Y >= 0; --Y) {
            const int xy = 6 * X + Y;
            ts.toff[xy] = cap;
            ts.tcap[xy] = 0;
            const int k0 = pl->cls_begin[Y], k1 = pl->cls_begin[Y + 1];
            for (int P = pl->cls_begin[X]; P < pl->cls_begin[X + 1] && k1 > k0; ++P) {
                int end = k1;
                if (thr > 0.0) {
                    int lo = k0, hi = k1;
                    while (lo < hi) {
                        const int mid = (lo + hi) / 2;
                        if (pl->Q[P] * pl->Q[mid] >= thr) lo = mid + 1;
                        else hi = mid;
                    }
                    end = lo;
                }
                if (X == Y) end = std::min(end, P + 1);
                for (int s0 = k0; s0 < end; s0 += kRouteSeg) {
                    const int s1 = std::min(end, s0 + kRouteSeg);
                    rows.push_back(make_int4(P, s0, s1, xy));
                    rbase.push_back(q);
                    rtask.push_back(cap + ts.tcap[xy]);
                    q += s1 - s0;
                    ts.tcap[xy] += (s1 - s0 + 63) / 64;
                }
            }
            cap += ts.tcap[xy];
        }
    if (q >= (1ll << 55)) throw std::runtime_error("tlno_gpu: quartet list too large for the task encoding");
    ts.nrows = static_cast<long long>(rows.size());
    ts.qtotal = q;
    ts.tcap_total = cap;
    dev_upload(ts.d_rows, rows);
    dev_upload(ts.d_rtask, rtask);
    rbase.push_back(q);  // never empty: marks the list as built
    dev_upload(ts.d_rbase, rbase);
    if (q > pl->qlist_cap) {
        if (pl->d_qlist) HIP_OK(hipFree(pl->d_qlist));
        pl->d_qlist = nullptr;
        HIP_OK(hipMalloc(&pl->d_qlist, q * sizeof(int)));
        pl->qlist_cap = q;
    }
    if (cap > pl->tlist_cap) {
        if (pl->d_t64) HIP_OK(hipFree(pl->d_t64));
        if (pl->d_t32) HIP_OK(hipFree(pl->d_t32));
        pl->d_t64 = pl->d_t32 = nullptr;
        HIP_OK(hipMalloc(&pl->d_t64, cap * sizeof(int4)));
        HIP_OK(hipMalloc(&pl->d_t32, cap * sizeof(int4)));
        pl->tlist_cap = cap;
    }
}

// Buffers of the learned build's two-pass routing: the density survivors of
// every row (same layout as qlist) and their count per row.
void ensure_route_net(tlno_gpu_plan* pl, TaskSet& ts) {
    if (!ts.d_rcnt) {
        require_alloc("learned-screen routing buffers");
        HIP_OK(hipMalloc(&ts.d_rcnt, std::max<long long>(1, ts.nrows) * sizeof(int)));
    }
    if (ts.qtotal > pl->cand_cap) {
        require_alloc("learned-screen routing buffers");
        if (pl->d_cand) HIP_OK(hipFree(pl->d_cand));
        pl->d_cand = nullptr;
        pl->cand_cap = 0;
        HIP_OK(hipMalloc(&pl->d_cand, ts.qtotal * sizeof(int)));
        pl->cand_cap = ts.qtotal;
    }
}

// bra primitive slices per list-mode task so that a small launch still
// spreads over the machine (same policy as the range-mode task split)
int list_split(long long ntask, int nchunk) {
    const long long waves = ntask * nchunk;
    constexpr long long kTargetWaves = TLNO_TARGET_WAVES;
    if (waves <= 0 || waves >= kTargetWaves) return 1;
    return static_cast<int>(std::min<long long>(256, (kTargetWaves + waves - 1) / waves));
}

// Estimated cost (ms) of a class pair's share of one build, for ordering its
// launches: the autotuner's measurement when there is one, else primitive
// quartets times a per-primitive work model (roots, 2D tables, products).
double class_cost(const tlno_gpu_plan* pl, const TaskSet& ts, int X, int Y) {
    if (pl->cls_cost[X][Y] > 0.0) return pl->cls_cost[X][Y];
    const int LAB = kClassLAB[X], LCD = kClassLAB[Y];
    const int nroot = (LAB + LCD) / 2 + 1;
    const double work = 100.0 + nroot * (15.0 * (LAB + 1) * (LCD + 1) + 3.0 * kClassNcart[X] * kClassNcart[Y]);
    return static_cast<double>(ts.cls_pq[X][Y]) * work * 1e-10;
}

// The team evaluator's quartet list of class pair (X, Y), expanded on the
// device from the range tasks on first use (one-off: allocates, then waits
// for the plan's stream so the staging buffer can go).
const int2* ensure_team_quartets(tlno_gpu_plan* pl, TaskSet& ts, int X, int Y) {
    if (ts.d_tq[X][Y]) return ts.d_tq[X][Y];
    const std::vector<long long>& cum = ts.cum_sq[X][Y];
    const long long nt = ts.cnt[X][Y];
    if (nt <= 0 || static_cast<long long>(cum.size()) != nt + 1 || cum.back() <= 0) return nullptr;
    long long* d_cum = nullptr;
    require_alloc("team-evaluator quartet lists");
    HIP_OK(hipMalloc(&ts.d_tq[X][Y], cum.back() * sizeof(int2)));
    HIP_OK(hipMalloc(&d_cum, cum.size() * sizeof(long long)));
    HIP_OK(hipMemcpyAsync(d_cum, cum.data(), cum.size() * sizeof(long long), hipMemcpyHostToDevice, pl->stream));
    k_team_quartets<<<static_cast<unsigned>(nt), 64, 0, pl->stream>>>(ts.d_tasks + ts.off[X][Y],
                                                                       static_cast<int>(nt), d_cum, ts.d_tq[X][Y]);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(pl->stream));
    HIP_OK(hipFree(d_cum));
    // heaviest quartets first (primitive quartets, stable): a wave lasts as
    // long as its quartet's combinations, which range over three orders of
    // magnitude where contraction lengths differ, and waves start in list
    // order, so the longest ones no longer trail at the end of the launch.
    // The team evaluator always takes a class's whole list (it keeps the
    // FP32 share of a mixed build too), so the order is free.
    std::vector<int2> h(static_cast<std::size_t>(cum.back()));
    HIP_OK(hipMemcpy(h.data(), ts.d_tq[X][Y], h.size() * sizeof(int2), hipMemcpyDeviceToHost));
    std::stable_sort(h.begin(), h.end(), [&](const int2& u, const int2& v) {
        return static_cast<long long>(pl->pn_host[u.x]) * pl->pn_host[u.y] >
               static_cast<long long>(pl->pn_host[v.x]) * pl->pn_host[v.y];
    });
    HIP_OK(hipMemcpy(ts.d_tq[X][Y], h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice));
    return ts.d_tq[X][Y];
}

// Work buffers of one build: the plan's own, or the device's static scratch
// (allocation-free builds).
struct BuildBuffers {
    double* Dint;
    double* Jacc;
    double* Kacc;
    unsigned long long* counter;
    double* scalar;
};

BuildBuffers plan_buffers(tlno_gpu_plan* pl) { return {pl->dDint, pl->dJacc, pl->dKacc, pl->d_counter, pl->d_scalar}; }

void jk_device(tlno_gpu_plan* pl, const double* dD, double* dJ, double* dK, double thr, const BuildBuffers& bb) {
    DeviceCtx& ctx = device_ctx(pl->device);
    HIP_OK(hipSetDevice(pl->device));
    if (!std::isfinite(thr) || thr < 0.0)
        throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
    TaskSet& ts = get_tasks(pl, thr);
    const hipStream_t st = pl->stream;
    const long mc2 = static_cast<long>(pl->Mcart) * pl->Mcart;
    const long mo2 = static_cast<long>(pl->Mout) * pl->Mout;
    const C2SArgs c2s{ctx.c2s};
    if (pl->spherical)
        k_prep_sph<<<grid_for(mc2, 256), 256, 0, st>>>(dD, bb.Dint, pl->d_intL, pl->d_intloc, pl->d_intsb,
                                                       pl->Mcart, pl->Msph, c2s);
    else
        k_prep_cart<<<grid_for(mc2, 256), 256, 0, st>>>(dD, bb.Dint, pl->d_int2basis, pl->Mcart);
    HIP_OK(hipGetLastError());
    if (dJ) k_fill<<<grid_for(mc2, 256), 256, 0, st>>>(bb.Jacc, mc2, 0.0);
    if (dK) k_fill<<<grid_for(mc2, 256), 256, 0, st>>>(bb.Kacc, mc2, 0.0);
    JKArgs a{};
    a.pd = pl->pair_data();
    a.tb = ctx.tb;
    a.D = bb.Dint;
    a.J = bb.Jacc;
    a.K = bb.Kacc;
    a.M = pl->Mcart;
    a.do_j = dJ != nullptr;
    a.do_k = dK != nullptr;
    double maxD = 0.0;
    if (pl->fp32_gate > 0.0 && (a.do_j || a.do_k)) {
        // max |D| in the caller's basis, the same basis as the Schwarz bounds
        k_fill<<<1, 64, 0, st>>>(bb.scalar, 1, 0.0);
        k_absmax<<<static_cast<unsigned>(std::min<long>(1024, grid_for(mo2, 256))), 256, 0, st>>>(
            dD, mo2, reinterpret_cast<unsigned long long*>(bb.scalar));
        HIP_OK(hipMemcpyAsync(&maxD, bb.scalar, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
    }
    pl->last_maxD = maxD;
    a.cut = nullptr;
    a.qlist = nullptr;
    a.counter = bb.counter;
    k_fill<<<grid_for(4 * kCounterSlots, 256), 256, 0, st>>>(reinterpret_cast<double*>(bb.counter), 4 * kCounterSlots, 0.0);
    if (pl->density_thr > 0.0 && (a.do_j || a.do_k)) {
        const int ng = static_cast<int>(pl->groups.size());
        k_fill<<<grid_for(ng, 256), 256, 0, st>>>(reinterpret_cast<double*>(pl->d_gmax), ng, 0.0);
        k_fill<<<1, 64, 0, st>>>(reinterpret_cast<double*>(pl->d_clsmax), 16, 0.0);
        k_group_rowmax<<<pl->Mout, 256, 0, st>>>(dD, pl->d_row_group, pl->Mout, pl->d_gmax);
        k_pair_factor<<<grid_for(pl->npair, 256), 256, 0, st>>>(pl->d_gmax, pl->d_pgA, pl->d_pgB, pl->d_pcls,
                                                                pl->npair, pl->d_pairfac, pl->d_clsmax);
        const int nlow = pl->cls_begin[kNumLowClasses];  // f pairs: per-quartet test in the team kernel
        if (nlow > 0)
            k_density_cut<<<grid_for(nlow, 256), 256, 0, st>>>(pl->d_Q, pl->d_pcls, pl->d_clsb, pl->d_pairfac,
                                                               pl->d_clsmax, nlow, pl->density_thr, pl->d_cut);
        const long long ng2 = static_cast<long long>(ng) * ng;
        if (ng2 > 0)
            k_block_max<<<grid_for(ng2, 256), 256, 0, st>>>(dD, pl->d_grows, ng, pl->Mout, pl->d_blk, pl->d_bsum);
        HIP_OK(hipGetLastError());
    }
    a.split = 1;
    if (pl->density_thr > 0.0 && pl->decay_thr > 0.0)
        throw std::runtime_error(
            "eri_ao_tensor_hgp: density screening requires Schwarz bounds and cannot be combined with the "
            "heuristic decay metric.");
    a.fs_min = pl->fastschwarz > 0.0
                   ? pl->fastschwarz * pl->fastschwarz / (kFastSchwarzPrefactor * kFastSchwarzPrefactor)
                   : 0.0;
    const bool route = (pl->density_thr > 0.0 || pl->learned || pl->decay_thr > 0.0) && (a.do_j || a.do_k);
    unsigned tcount[72] = {};
    if (route) {
        ensure_route(pl, ts, thr);
        k_fill<<<grid_for(4 * kCounterSlots, 256), 256, 0, st>>>(reinterpret_cast<double*>(pl->d_rstats),
                                                                4 * kCounterSlots, 0.0);
        RouteArgs ra;
        ra.rows = ts.d_rows;
        ra.rbase = ts.d_rbase;
        ra.nrows = ts.nrows;
        ra.cut = pl->density_thr > 0.0 ? pl->d_cut : nullptr;
        ra.Q = pl->d_Q;
        ra.pgA = pl->d_pgA;
        ra.pgB = pl->d_pgB;
        ra.ng = static_cast<int>(pl->groups.size());
        ra.blk = pl->density_thr > 0.0 ? pl->d_blk : nullptr;
        ra.bsum = pl->d_bsum;
        ra.mode = (a.do_j && a.do_k) ? 0 : (a.do_j ? 1 : 2);
        ra.tau = pl->density_thr;
        ra.thr = thr;
        ra.mlp = MlpPacked{pl->d_frag, pl->d_mvec, pl->d_mnorm, pl->mb3};
        ra.fa = FeatArgs{pl->d_pf0, pl->d_pf1, pl->d_pcen};
        ra.margin = pl->learned_margin;
        ra.decay = pl->decay_thr > 0.0 ? pl->d_decay : nullptr;
        ra.decay_thr = pl->decay_thr;
        ra.fp32_cls = 0;
        if (pl->fp32_gate > 0.0)
            for (int X = 0; X < 6; ++X)
                for (int Y = 0; Y <= X; ++Y)
                    if (!pl->fp32_slower[X][Y] && !use_team(g_kernels[X][Y], false, a, class_team_min(pl, X, Y)))
                        ra.fp32_cls |= 1ull << (6 * X + Y);
        ra.gate = pl->fp32_gate;
        ra.maxD = bb.scalar;
        ra.qlist = pl->d_qlist;
        ra.t64 = pl->d_t64;
        ra.t32 = pl->d_t32;
        ra.rtask = ts.d_rtask;
        ra.stats = pl->d_rstats;
        ra.cand = nullptr;
        ra.rcnt = nullptr;
        if (ts.nrows > 0) {
            const unsigned grid = grid_for(ts.nrows * 64, 256);
            if (pl->learned) {
                ensure_route_net(pl, ts);
                ra.cand = pl->d_cand;
                ra.rcnt = ts.d_rcnt;
                k_route<true><<<grid, 256, 0, st>>>(ra);
                k_route_net<<<grid, 256, 0, st>>>(ra);
            } else {
                k_route<false><<<grid, 256, 0, st>>>(ra);
            }
        }
        HIP_OK(hipGetLastError());
        // every row's tasks sit in its fixed slots: each class launch covers
        // the class's slot range (empty tasks leave at once)
        for (int xy = 0; xy < 36; ++xy) {
            tcount[xy] = static_cast<unsigned>(ts.tcap[xy]);
            tcount[36 + xy] = ((ra.fp32_cls >> xy) & 1ull) ? static_cast<unsigned>(ts.tcap[xy]) : 0u;
        }
    }
    if (a.do_j || a.do_k) {
        // Class-pair kernels are independent (their J/K contributions are
        // atomic adds), so they run concurrently on the plan's side streams,
        // heaviest first; timing mode keeps them on one stream so each
        // class's events bracket only its own kernel.
        const bool concurrent = !pl->timing;
        std::unique_lock<std::mutex> side_lk;
        if (concurrent) {
            side_lk = std::unique_lock<std::mutex>(ctx.side_mu);
            ensure_side_streams(ctx);
            HIP_OK(hipEventRecord(ctx.ev_fork, st));
            for (int s = 0; s < ctx.nside; ++s) HIP_OK(hipStreamWaitEvent(ctx.side[s], ctx.ev_fork, 0));
        }
        // Every class launch (FP64 share, FP32 share) becomes a job with an
        // estimated cost; jobs are collected heaviest class first, then
        // launched.
        struct Job {
            int X, Y, team_min;
            bool fp32;
            JKArgs args;
            long long nwaves;
            double cost;
            bool high;  // class pair with an f shell (team kernel of g_high)
        };
        // at most two jobs (FP64 and FP32 share) per class pair of shells up
        // to d, one per class pair with an f shell; a fixed array keeps the
        // host side free of heap allocations as well
        constexpr std::size_t kMaxJobs = 72 + kNumHigh * kNumClasses;
        Job jobs[kMaxJobs];
        std::size_t njobs = 0;
        auto add_jobs = [&](int X, int Y, bool fp32, const JKArgs& args, long long nwaves, int team_min,
                            double cost) {
            // the FP32 share of a class whose FP32 kernels measured slower
            // runs the FP64 kernels on the same tasks
            if (fp32 && pl->fp32_slower[X][Y]) fp32 = false;
            if (njobs == kMaxJobs) throw std::runtime_error("tlno_gpu: too many class launches");
            jobs[njobs++] = Job{X, Y, team_min, fp32, args, nwaves, cost, false};
        };
        for (int X = 5; X >= 0 && route; --X)
            for (int Y = X; Y >= 0; --Y) {
                const int xy = 6 * X + Y;
                const long long n64 = tcount[xy], n32 = tcount[36 + xy];
                if (n64 + n32 == 0) continue;
                const ClassKernels& k = g_kernels[X][Y];
                const double c = class_cost(pl, ts, X, Y) / static_cast<double>(n64 + n32);
                a.cut = nullptr;
                a.qlist = pl->d_qlist;
                if (n64 > 0) {
                    a.tasks = pl->d_t64 + ts.toff[xy];
                    a.ntasks = static_cast<int>(n64);
                    a.split = use_team(k, false, a, class_team_min(pl, X, Y)) ? 1 : list_split(n64, k.nchunk);
                    a.team_slices = team_slices_for(ts.cls_sq[X][Y]);
                    add_jobs(X, Y, false, a, n64 * a.split, class_team_min(pl, X, Y), c * n64);
                }
                if (n32 > 0) {
                    a.tasks = pl->d_t32 + ts.toff[xy];
                    a.ntasks = static_cast<int>(n32);
                    if (use_team(k, false, a, class_team_min(pl, X, Y))) {
                        // the team evaluator (FP64) outruns the FP32 lane
                        // evaluator on its classes: the FP32 share stays on it
                        a.split = 1;
                        a.team_slices = team_slices_for(ts.cls_sq[X][Y]);
                        add_jobs(X, Y, false, a, n32, class_team_min(pl, X, Y), c * n32);
                    } else {
                        a.split = list_split(n32, k.nchunk);
                        add_jobs(X, Y, true, a, n32 * a.split, 0, c * n32);
                    }
                }
            }
        a.qlist = nullptr;
        a.split = 1;
        for (int X = 5; X >= 0 && !route; --X)
            for (int Y = X; Y >= 0; --Y) {
                if (ts.lcnt[X][Y] > 0 && !use_team(g_kernels[X][Y], false, a, class_team_min(pl, X, Y))) {
                    // lane evaluator: bucketed list tasks (kets of one task have
                    // similar primitive counts); the team evaluator (one wave per
                    // quartet) keeps the range tasks
                    const long long nt = ts.lcnt[X][Y];
                    long long n64 = nt;
                    if (pl->fp32_gate > 0.0) {
                        const auto& bm = ts.lbmax[X][Y];
                        const double cut = pl->fp32_gate / std::max(maxD, 1e-300);
                        n64 = std::lower_bound(bm.begin(), bm.end(), cut,
                                               [](double v, double c) { return v >= c; }) - bm.begin();
                    }
                    const ClassKernels& k = g_kernels[X][Y];
                    const double c = class_cost(pl, ts, X, Y) / static_cast<double>(nt);
                    JKArgs al = a;
                    al.cut = nullptr;
                    al.qlist = ts.d_bk;
                    if (n64 > 0) {
                        // a whole FP64 list runs heaviest slices first (the
                        // launch's tail is its last-started waves)
                        al.tasks = (n64 == nt ? ts.d_ltasks_lpt : ts.d_ltasks) + ts.loff[X][Y];
                        al.ntasks = static_cast<int>(n64);
                        al.split = 0;  // the tasks carry their cost-balanced slices
                        add_jobs(X, Y, false, al, n64, class_team_min(pl, X, Y), c * n64);
                    }
                    if (nt > n64) {
                        al.tasks = (n64 == 0 ? ts.d_ltasks_lpt : ts.d_ltasks + n64) + ts.loff[X][Y];
                        al.ntasks = static_cast<int>(nt - n64);
                        al.split = 0;
                        add_jobs(X, Y, true, al, nt - n64, 0, c * (nt - n64));
                    }
                    continue;
                }
                const long long nt = ts.cnt[X][Y];
                if (nt == 0) continue;
                // FP64 prefix / FP32 suffix of the bound-ordered task list (a
                // class on the team evaluator, FP64 and faster there than the
                // FP32 lane evaluator, keeps all its quartets)
                long long n64 = nt;
                if (pl->fp32_gate > 0.0 && !use_team(g_kernels[X][Y], false, a, class_team_min(pl, X, Y))) {
                    const auto& bm = ts.bmax[X][Y];
                    const double cut = pl->fp32_gate / std::max(maxD, 1e-300);
                    n64 = std::lower_bound(bm.begin(), bm.end(), cut, [](double v, double c) { return v >= c; }) -
                          bm.begin();
                }
                const double c = class_cost(pl, ts, X, Y) / static_cast<double>(nt);
                a.cut = nullptr;  // density screening always takes the routed path
                if (n64 > 0) {
                    a.tasks = ts.d_tasks + ts.off[X][Y];
                    a.ntasks = static_cast<int>(n64);
                    a.team_slices = team_slices_for(ts.cls_sq[X][Y]);
                    // the team evaluator takes one wave per listed quartet
                    // (the range tasks hold empty ket slots and the lane
                    // evaluator's bra primitive slices)
                    if (use_team(g_kernels[X][Y], false, a, class_team_min(pl, X, Y))) {
                        a.tquart = ensure_team_quartets(pl, ts, X, Y);
                        a.ntquart = a.tquart ? ts.cum_sq[X][Y][n64] : 0;
                    }
                    add_jobs(X, Y, false, a, n64, class_team_min(pl, X, Y), c * n64);
                    a.tquart = nullptr;
                    a.ntquart = 0;
                }
                if (nt > n64) {
                    a.tasks = ts.d_tasks + ts.off[X][Y] + n64;
                    a.ntasks = static_cast<int>(nt - n64);
                    add_jobs(X, Y, true, a, nt - n64, 0, c * (nt - n64));
                }
            }
        // Class pairs with an f shell: the team evaluator over the plan's
        // quartet lists, in FP64, with the density and decay screens applied
        // per quartet inside the kernel (the routing kernel and the learned
        // screen's network cover shells up to d; these quartets take the
        // density rule alone)
        for (int X = kNumClasses - 1; X >= kNumLowClasses; --X)
            for (int Y = X; Y >= 0; --Y) {
                const int h = X - kNumLowClasses;
                const long long n = ts.hcnt[h][Y];
                if (n == 0) continue;
                JKArgs ah = a;
                ah.tasks = nullptr;
                ah.ntasks = 0;
                ah.qlist = nullptr;
                ah.cut = nullptr;
                ah.split = 1;
                ah.tquart = ts.d_hq + ts.hoff[h][Y];
                ah.ntquart = n;
                ah.team_slices = team_slices_for(n);
                if (pl->density_thr > 0.0) {
                    ah.sblk = pl->d_blk;
                    ah.spgA = pl->d_pgA;
                    ah.spgB = pl->d_pgB;
                    ah.sQ = pl->d_Q;
                    ah.sng = static_cast<int>(pl->groups.size());
                    ah.smode = (a.do_j && a.do_k) ? 0 : (a.do_j ? 1 : 2);
                    ah.stau = pl->density_thr;
                }
                if (pl->decay_thr > 0.0) {
                    ah.sdecay = pl->d_decay;
                    ah.sdecay_thr = pl->decay_thr;
                }
                if (njobs == kMaxJobs) throw std::runtime_error("tlno_gpu: too many class launches");
                const double cost = 1e-6 * static_cast<double>(n) * kClassNcart[X] * kClassNcart[Y];
                ensure_gwork(pl, g_high[h][Y].gwork);
                jobs[njobs++] = Job{X, Y, 1, false, ah, n, cost, true};
            }
        if (concurrent) {
            // each job onto the side stream with the least estimated work so
            // far (list scheduling in class order; ordering by cost or
            // splitting the chunk groups of a class over streams measured no
            // better: concurrent kernels do not add up their serial times)
            double load[kSideStreams] = {};
            for (std::size_t i = 0; i < njobs; ++i) {
                Job& j = jobs[i];
                int best = 0;
                for (int q = 1; q < ctx.nside; ++q)
                    if (load[q] < load[best]) best = q;
                // the g classes with global working sets share the plan's
                // scratch: all on the first side stream, one after another
                if (j.high && g_high[j.X - kNumLowClasses][j.Y].gwork > 0) best = 0;
                load[best] += j.cost;
                if (j.high)
                    launch_high(g_high[j.X - kNumLowClasses][j.Y], j.args, ctx.side[best], pl->d_gwork, pl->gslots);
                else
                    launch_jk(g_kernels[j.X][j.Y], j.fp32, j.args, j.nwaves, ctx.side[best], j.team_min);
            }
        } else {
            // timing mode: one stream, class order, events around each class's jobs
            for (std::size_t i = 0; i < njobs; ++i) {
                Job& j = jobs[i];
                if (j.high) {
                    hipEvent_t* e = pl->hev[j.X - kNumLowClasses][j.Y];
                    if (e[0]) HIP_OK(hipEventRecord(e[0], st));
                    launch_high(g_high[j.X - kNumLowClasses][j.Y], j.args, st, pl->d_gwork, pl->gslots);
                    if (e[1]) HIP_OK(hipEventRecord(e[1], st));
                    continue;
                }
                const bool first = i == 0 || jobs[i - 1].X != j.X || jobs[i - 1].Y != j.Y;
                const bool last = i + 1 == njobs || jobs[i + 1].X != j.X || jobs[i + 1].Y != j.Y;
                if (first) HIP_OK(hipEventRecord(pl->ev[j.X][j.Y][0], st));
                launch_jk(g_kernels[j.X][j.Y], j.fp32, j.args, j.nwaves, st, j.team_min);
                if (last) HIP_OK(hipEventRecord(pl->ev[j.X][j.Y][1], st));
            }
        }
        if (concurrent)
            for (int s = 0; s < ctx.nside; ++s) {
                HIP_OK(hipEventRecord(ctx.ev_join[s], ctx.side[s]));
                HIP_OK(hipStreamWaitEvent(st, ctx.ev_join[s], 0));
            }
    }
    if (dJ) {
        if (pl->spherical)
            k_fin_sph<<<grid_for(mo2, 256), 256, 0, st>>>(bb.Jacc, dJ, pl->d_sphL, pl->d_sphloc, pl->d_sphib,
                                                          pl->Mcart, pl->Msph, c2s);
        else
            k_fin_cart<<<grid_for(mo2, 256), 256, 0, st>>>(bb.Jacc, dJ, pl->d_int2basis, pl->Mcart);
    }
    if (dK) {
        if (pl->spherical)
            k_fin_sph<<<grid_for(mo2, 256), 256, 0, st>>>(bb.Kacc, dK, pl->d_sphL, pl->d_sphloc, pl->d_sphib,
                                                          pl->Mcart, pl->Msph, c2s);
        else
            k_fin_cart<<<grid_for(mo2, 256), 256, 0, st>>>(bb.Kacc, dK, pl->d_int2basis, pl->Mcart);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
    {
        unsigned long long evaluated[4] = {0, 0, 0, 0};
        unsigned long long slots[4 * kCounterSlots];
        HIP_OK(hipMemcpy(slots, bb.counter, sizeof(slots), hipMemcpyDeviceToHost));
        for (int i = 0; i < 4 * kCounterSlots; ++i) evaluated[i % 4] += slots[i];
        pl->last_evaluated = static_cast<long long>(evaluated[0]);
        pl->last_evaluated_prim = static_cast<long long>(evaluated[1]);
        pl->last_fp32_sq = static_cast<long long>(evaluated[2]);
        pl->last_fp32_pq = static_cast<long long>(evaluated[3]);
        if (route) {
            unsigned long long rs[4 * kCounterSlots];
            HIP_OK(hipMemcpy(rs, pl->d_rstats, sizeof(rs), hipMemcpyDeviceToHost));
            for (int i = 0; i < 4; ++i) pl->last_route[i] = 0;
            for (int i = 0; i < 4 * kCounterSlots; ++i) pl->last_route[i % 4] += static_cast<long long>(rs[i]);
        } else {
            for (int i = 0; i < 4; ++i) pl->last_route[i] = 0;
        }
    }
    pl->last_shell_quartets = ts.shell_quartets;
    pl->last_prim_quartets = ts.prim_quartets;
    pl->last_thr = thr;
    if (pl->timing && (a.do_j || a.do_k)) {
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y <= X; ++Y)
                if (ts.cnt[X][Y]) {
                    float ms = 0.0f;
                    HIP_OK(hipEventElapsedTime(&ms, pl->ev[X][Y][0], pl->ev[X][Y][1]));
                    pl->cls_ms[X][Y] += ms;
                }
        for (int h = 0; h < kNumHigh; ++h)
            for (int Y = 0; Y <= h + kNumLowClasses; ++Y)
                if (ts.hcnt[h][Y] && pl->hev[h][Y][0]) {
                    float ms = 0.0f;
                    HIP_OK(hipEventElapsedTime(&ms, pl->hev[h][Y][0], pl->hev[h][Y][1]));
                    pl->hcls_ms[h][Y] += ms;
                }
        ++pl->timed_builds;
    }
}

// ---- canonical ERI stream checksum ----
// Sum of every unique spherical (ab|cd), a >= b, c >= d, (ab) >= (cd) (pair
// index), over Schwarz-surviving shell quartets, the value the reference's
// packed endpoint returns (rys_engine.ipp:270-300, 17096-17103).  Shell
// quartets with P != Q contribute sum_abcd wP[ab] wQ[cd] (ab|cd)_cart with
//   wP[a][b] = u[a] u[b], u[a] = sum_s C[s][a]          (A != B),
//   wP[a][b] = sum_{s >= t} C[s][a] C[t][b]             (A == B),
// evaluated inside the J/K kernels (checksum mode); the npair quartets (P|P)
// are projected and summed on the host with the (ab) >= (cd) rule.
double eri_checksum(tlno_gpu_plan* pl, double thr) {
    if (!pl->spherical)
        throw std::runtime_error("eri_ao_tensor_hgp: the ERI stream is pure spherical; build a spherical plan.");
    if (!std::isfinite(thr) || thr < 0.0)
        throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
    DeviceCtx& ctx = device_ctx(pl->device);
    HIP_OK(hipSetDevice(pl->device));
    const C2S& cs = c2s_table();
    // weights per pair kind (2 * class + diag)
    std::vector<double> cw(12 * 36, 0.0);
    for (int c = 0; c < 6; ++c) {
        const int LA = (c == 0 ? 0 : (c <= 2 ? 1 : 2)), LB = (c == 0 || c == 1 || c == 3) ? 0 : (c == 2 || c == 4 ? 1 : 2);
        const int na = ncart(LA), nb = ncart(LB);
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) {
                double ua = 0.0, ub = 0.0;
                for (int t = 0; t < 2 * LA + 1; ++t) ua += cs.c[LA][t][a];
                for (int t = 0; t < 2 * LB + 1; ++t) ub += cs.c[LB][t][b];
                cw[(2 * c) * 36 + a * nb + b] = ua * ub;
                if (LA == LB) {
                    double w = 0.0;
                    for (int s1 = 0; s1 < 2 * LA + 1; ++s1)
                        for (int t1 = 0; t1 <= s1; ++t1) w += cs.c[LA][s1][a] * cs.c[LB][t1][b];
                    cw[(2 * c + 1) * 36 + a * nb + b] = w;
                }
            }
    }
    double* dcw = nullptr;
    double* dchk = nullptr;
    HIP_OK(hipMalloc(&dcw, cw.size() * sizeof(double)));
    HIP_OK(hipMalloc(&dchk, sizeof(double)));
    HIP_OK(hipMemcpy(dcw, cw.data(), cw.size() * sizeof(double), hipMemcpyHostToDevice));
    const hipStream_t st = pl->stream;
    k_fill<<<1, 64, 0, st>>>(dchk, 1, 0.0);
    const TaskSet& ts = get_tasks(pl, thr);
    JKArgs a{};
    a.pd = pl->pair_data();
    a.tb = ctx.tb;
    a.M = pl->Mcart;
    a.split = 1;
    a.cw = dcw;
    a.chk = dchk;
    for (int X = 5; X >= 0; --X)
        for (int Y = X; Y >= 0; --Y) {
            const long long nt = ts.cnt[X][Y];
            if (nt == 0) continue;
            a.tasks = ts.d_tasks + ts.off[X][Y];
            a.ntasks = static_cast<int>(nt);
            launch_jk(g_kernels[X][Y], false, a, nt, st);
        }
    HIP_OK(hipGetLastError());
    double total = 0.0;
    HIP_OK(hipMemcpyAsync(&total, dchk, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    HIP_OK(hipFree(dcw));
    HIP_OK(hipFree(dchk));
    // (P|P) quartets: full spherical blocks, (ab) >= (cd) and a >= b, c >= d
    auto pair_index = [](long long i, long long j) {
        if (i < j) std::swap(i, j);
        return i * (i + 1) / 2 + j;
    };
    for (int X = 0; X < kNumClasses; ++X) {
        const int b0 = pl->cls_begin[X], b1 = pl->cls_begin[X + 1];
        std::vector<int2> qs;
        std::vector<long long> off;
        const int LA = kClassL[X][0], LB = kClassL[X][1];
        const int na = ncart(LA), nb = ncart(LB), nab = na * nb;
        for (int P = b0; P < b1; ++P)
            if (pl->Q[P] * pl->Q[P] >= thr) {
                qs.push_back(make_int2(P, P));
                off.push_back(static_cast<long long>(qs.size() - 1) * nab * nab);
            }
        if (qs.empty()) continue;
        double* dout = nullptr;
        HIP_OK(hipMalloc(&dout, qs.size() * nab * nab * sizeof(double)));
        run_store(pl, X, X, qs, off, dout);
        std::vector<double> h(qs.size() * nab * nab);
        HIP_OK(hipMemcpy(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIP_OK(hipFree(dout));
        const int sa = 2 * LA + 1, sb = 2 * LB + 1;
        std::vector<double> half(static_cast<std::size_t>(sa) * sb * nab), sph(static_cast<std::size_t>(sa) * sb * sa * sb);
        for (std::size_t k = 0; k < qs.size(); ++k) {
            const int P = qs[k].x;
            const GroupH& Ga = pl->groups[pl->pgA[P]];
            const GroupH& Gb = pl->groups[pl->pgB[P]];
            const double* blk = h.data() + k * nab * nab;
            // bra projection, then ket projection
            for (int s1 = 0; s1 < sa; ++s1)
                for (int t1 = 0; t1 < sb; ++t1)
                    for (int cd = 0; cd < nab; ++cd) {
                        double v = 0.0;
                        for (int a1 = 0; a1 < na; ++a1)
                            for (int b1i = 0; b1i < nb; ++b1i)
                                v += cs.c[LA][s1][a1] * cs.c[LB][t1][b1i] * blk[(a1 * nb + b1i) * nab + cd];
                        half[(static_cast<std::size_t>(s1) * sb + t1) * nab + cd] = v;
                    }
            for (int st1 = 0; st1 < sa * sb; ++st1)
                for (int s2 = 0; s2 < sa; ++s2)
                    for (int t2 = 0; t2 < sb; ++t2) {
                        double v = 0.0;
                        for (int c1 = 0; c1 < na; ++c1)
                            for (int d1 = 0; d1 < nb; ++d1)
                                v += cs.c[LA][s2][c1] * cs.c[LB][t2][d1] * half[static_cast<std::size_t>(st1) * nab + c1 * nb + d1];
                        sph[static_cast<std::size_t>(st1) * sa * sb + s2 * sb + t2] = v;
                    }
            const bool diag = pl->pgA[P] == pl->pgB[P];
            for (int s1 = 0; s1 < sa; ++s1)
                for (int t1 = 0; t1 < sb; ++t1) {
                    const long long i1 = Ga.sph_base + s1, j1 = Gb.sph_base + t1;
                    if (diag && j1 > i1) continue;
                    const long long p1 = pair_index(i1, j1);
                    for (int s2 = 0; s2 < sa; ++s2)
                        for (int t2 = 0; t2 < sb; ++t2) {
                            const long long i2 = Ga.sph_base + s2, j2 = Gb.sph_base + t2;
                            if (diag && j2 > i2) continue;
                            if (pair_index(i2, j2) > p1) continue;
                            total += sph[(static_cast<std::size_t>(s1) * sb + t1) * sa * sb + s2 * sb + t2];
                        }
                }
        }
    }
    // quartets with an f shell and P != Q: Cartesian blocks from the team
    // kernel, folded on the host with the same pair weights
    auto pair_weights = [&](int P, std::vector<double>& w) {
        const int LA = pl->groups[pl->pgA[P]].L, LB = pl->groups[pl->pgB[P]].L;
        const int na = ncart(LA), nb = ncart(LB);
        const bool same = pl->pgA[P] == pl->pgB[P];
        w.assign(static_cast<std::size_t>(na) * nb, 0.0);
        for (int a1 = 0; a1 < na; ++a1)
            for (int b1 = 0; b1 < nb; ++b1) {
                double v = 0.0;
                if (!same) {
                    double ua = 0.0, ub = 0.0;
                    for (int t = 0; t < 2 * LA + 1; ++t) ua += cs.c[LA][t][a1];
                    for (int t = 0; t < 2 * LB + 1; ++t) ub += cs.c[LB][t][b1];
                    v = ua * ub;
                } else {
                    for (int s1 = 0; s1 < 2 * LA + 1; ++s1)
                        for (int t1 = 0; t1 <= s1; ++t1) v += cs.c[LA][s1][a1] * cs.c[LB][t1][b1];
                }
                w[static_cast<std::size_t>(a1) * nb + b1] = v;
            }
    };
    for (int X = kNumLowClasses; X < kNumClasses; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const int nab = kClassNcart[X], ncd = kClassNcart[Y];
            std::vector<int2> all;
            for (int P = pl->cls_begin[X]; P < pl->cls_begin[X + 1]; ++P)
                for (int Q = pl->cls_begin[Y]; Q < pl->cls_begin[Y + 1] && Q < P; ++Q)
                    if (pl->Q[P] * pl->Q[Q] >= thr) all.push_back(make_int2(P, Q));
            constexpr std::size_t kBatchDoubles = std::size_t(1) << 24;
            const std::size_t per = static_cast<std::size_t>(nab) * ncd;
            const std::size_t nbatch = std::max<std::size_t>(1, kBatchDoubles / per);
            std::vector<double> wp, wq;
            for (std::size_t q0 = 0; q0 < all.size(); q0 += nbatch) {
                const std::size_t q1 = std::min(all.size(), q0 + nbatch);
                std::vector<int2> qs(all.begin() + q0, all.begin() + q1);
                std::vector<long long> off(qs.size());
                for (std::size_t k = 0; k < qs.size(); ++k) off[k] = static_cast<long long>(k * per);
                double* dout = nullptr;
                HIP_OK(hipMalloc(&dout, qs.size() * per * sizeof(double)));
                run_store(pl, X, Y, qs, off, dout);
                std::vector<double> h(qs.size() * per);
                HIP_OK(hipMemcpy(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost));
                HIP_OK(hipFree(dout));
                for (std::size_t k = 0; k < qs.size(); ++k) {
                    pair_weights(qs[k].x, wp);
                    pair_weights(qs[k].y, wq);
                    const double* blk = h.data() + k * per;
                    for (int ab = 0; ab < nab; ++ab) {
                        double v = 0.0;
                        for (int cd = 0; cd < ncd; ++cd) v += wq[cd] * blk[static_cast<std::size_t>(ab) * ncd + cd];
                        total += wp[ab] * v;
                    }
                }
            }
        }
    return total;
}

// ------------------------------------------------ stream fingerprint ---
// The reference's indexed fingerprint of the canonical spherical stream
// (rys_engine.ipp:87-105 weights, 270-300 accumulation; the values its
// validation build returns, native_quartets.ipp:759-765): over every unique
// (ab|cd) x with canonical index i = qidx(pair(a,b), pair(c,d)),
//   sum x, sum |x|, sum x^2, sum x w_k(i) (k = 1..3), max |x|, count,
// w_k(i) = (bits 40..63 of splitmix64(i ^ seed_k) - 2^23) / 2^23.
// Cartesian blocks of a batch of shell quartets are written by the block
// kernel, then one wave per quartet projects its block to spherical in LDS,
// applies the unique-index rule and adds its sums into 64 slots (slot =
// quartet mod 64; the host sums the slots in long double).
constexpr unsigned long long kFpSeed[3] = {0x243f6a8885a308d3ull, 0x13198a2e03707344ull, 0xa4093822299f31d0ull};

struct C2SSparse {
    int n[3][5];
    int idx[3][5][6];
    double c[3][5][6];
};

struct FPArgs {
    const double* blocks;  // nblk doubles per quartet, pair orientation
    const int2* quartets;
    int n;
    int LA, LB, LC, LD;
    const int4* pinfo;     // per pair {spherical base of A, of B, same shell, 0}
    C2SSparse cs;
    double* acc;                  // 64 slots x 6 sums
    unsigned long long* maxbits;  // 64 slots (bit patterns of non-negative doubles)
    unsigned long long* count;    // 64 slots
};

__device__ __forceinline__ unsigned long long fp_mix(unsigned long long v) {
    v += 0x9e3779b97f4a7c15ull;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

__device__ __forceinline__ double fp_weight(unsigned long long i, unsigned long long seed) {
    const long long m = static_cast<long long>((fp_mix(i ^ seed) >> 40) & 0xffffffull) - 0x800000ll;
    return static_cast<double>(m) * (1.0 / 8388608.0);
}

__device__ __forceinline__ long long canon_pair(long long i, long long j) {
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

__global__ __launch_bounds__(64) void k_fingerprint(FPArgs a) {
    __shared__ double sblk[1296];
    __shared__ double shalf[900];
    __shared__ double ssph[625];
    const int q = blockIdx.x;
    if (q >= a.n) return;
    const int lane = threadIdx.x;
    const int na = ncart(a.LA), nb = ncart(a.LB), nc = ncart(a.LC), nd = ncart(a.LD);
    const int ncd = nc * nd, nblk = na * nb * ncd;
    const int sa = 2 * a.LA + 1, sb = 2 * a.LB + 1, sc = 2 * a.LC + 1, sd = 2 * a.LD + 1;
    const double* blk = a.blocks + static_cast<long long>(q) * nblk;
    for (int i = lane; i < nblk; i += 64) sblk[i] = blk[i];
    __syncthreads();
    for (int i = lane; i < sa * sb * ncd; i += 64) {
        const int s = i / (sb * ncd), t = (i / ncd) % sb, cd = i % ncd;
        double v = 0.0;
        for (int k1 = 0; k1 < a.cs.n[a.LA][s]; ++k1)
            for (int k2 = 0; k2 < a.cs.n[a.LB][t]; ++k2)
                v += a.cs.c[a.LA][s][k1] * a.cs.c[a.LB][t][k2] *
                     sblk[(a.cs.idx[a.LA][s][k1] * nb + a.cs.idx[a.LB][t][k2]) * ncd + cd];
        shalf[i] = v;
    }
    __syncthreads();
    const int nsph = sa * sb * sc * sd;
    for (int i = lane; i < nsph; i += 64) {
        const int st = i / (sc * sd), u = (i / sd) % sc, w = i % sd;
        double v = 0.0;
        for (int k1 = 0; k1 < a.cs.n[a.LC][u]; ++k1)
            for (int k2 = 0; k2 < a.cs.n[a.LD][w]; ++k2)
                v += a.cs.c[a.LC][u][k1] * a.cs.c[a.LD][w][k2] *
                     shalf[st * ncd + a.cs.idx[a.LC][u][k1] * nd + a.cs.idx[a.LD][w][k2]];
        ssph[i] = v;
    }
    __syncthreads();
    const int2 pq = a.quartets[q];
    const int4 iP = a.pinfo[pq.x], iQ = a.pinfo[pq.y];
    const bool same_pair = pq.x == pq.y;
    double sum = 0.0, sabs = 0.0, ssq = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, mx = 0.0;
    int cnt = 0;
    for (int i = lane; i < nsph; i += 64) {
        const int s = i / (sb * sc * sd), t = (i / (sc * sd)) % sb, u = (i / sd) % sc, w = i % sd;
        const long long ia = iP.x + s, ib = iP.y + t, ic = iQ.x + u, id = iQ.y + w;
        if (iP.z && ib > ia) continue;
        if (iQ.z && id > ic) continue;
        const long long p1 = canon_pair(ia, ib), p2 = canon_pair(ic, id);
        if (same_pair && p2 > p1) continue;
        const unsigned long long idx = static_cast<unsigned long long>(canon_pair(p1, p2));
        const double x = ssph[i];
        const double ax = fabs(x);
        sum += x;
        sabs += ax;
        ssq += x * x;
        m1 += x * fp_weight(idx, kFpSeed[0]);
        m2 += x * fp_weight(idx, kFpSeed[1]);
        m3 += x * fp_weight(idx, kFpSeed[2]);
        mx = fmax(mx, ax);
        ++cnt;
    }
    sum = wave_sum(sum);
    sabs = wave_sum(sabs);
    ssq = wave_sum(ssq);
    m1 = wave_sum(m1);
    m2 = wave_sum(m2);
    m3 = wave_sum(m3);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, m, 64));
        cnt += __shfl_xor(cnt, m, 64);
    }
    if (lane == 0) {
        double* slot = a.acc + 6 * (q & 63);
        atomicAdd(slot + 0, sum);
        atomicAdd(slot + 1, sabs);
        atomicAdd(slot + 2, ssq);
        atomicAdd(slot + 3, m1);
        atomicAdd(slot + 4, m2);
        atomicAdd(slot + 5, m3);
        atomicMax(a.maxbits + (q & 63), static_cast<unsigned long long>(__double_as_longlong(mx)));
        atomicAdd(a.count + (q & 63), static_cast<unsigned long long>(cnt));
    }
}

// out = {checksum, sum |x|, sum x^2, m1, m2, m3, max |x|}, count = number of
// canonical values, over shell quartets with Q_ab Q_cd >= thr.
void eri_fingerprint(tlno_gpu_plan* pl, double thr, double* out, unsigned long long* count) {
    if (!pl->spherical)
        throw std::runtime_error("eri_ao_tensor_hgp: the ERI stream is pure spherical; build a spherical plan.");
    if (pl->cls_begin[kNumLowClasses] != pl->npair)
        throw std::runtime_error(
            "tlno_gpu: the stream fingerprint covers shells up to d in this build (the checksum covers f)");
    if (!std::isfinite(thr) || thr < 0.0)
        throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
    require_alloc("the stream fingerprint");
    DeviceCtx& ctx = device_ctx(pl->device);
    HIP_OK(hipSetDevice(pl->device));
    const hipStream_t st = pl->stream;
    FPArgs f{};
    const C2S& c2 = c2s_table();
    for (int L = 0; L < 3; ++L)
        for (int s = 0; s < 2 * L + 1; ++s) {
            int k = 0;
            for (int c = 0; c < ncart(L); ++c)
                if (c2.c[L][s][c] != 0.0) {
                    f.cs.idx[L][s][k] = c;
                    f.cs.c[L][s][k] = c2.c[L][s][c];
                    ++k;
                }
            f.cs.n[L][s] = k;
        }
    std::vector<int4> pinfo(static_cast<std::size_t>(pl->npair));
    for (int P = 0; P < pl->npair; ++P)
        pinfo[P] = make_int4(pl->groups[pl->pgA[P]].sph_base, pl->groups[pl->pgB[P]].sph_base,
                             pl->pgA[P] == pl->pgB[P] ? 1 : 0, 0);
    int4* dpi = nullptr;
    double* dacc = nullptr;
    unsigned long long *dmax = nullptr, *dcnt = nullptr;
    double* dblk = nullptr;
    constexpr long long kBatchDoubles = 32ll << 20;  // 256 MiB of blocks per launch pair
    struct Freer {
        std::vector<void*> p;
        ~Freer() {
            for (void* x : p) (void)hipFree(x);
        }
    } fr;
    HIP_OK(hipMalloc(&dpi, pinfo.size() * sizeof(int4)));
    fr.p.push_back(dpi);
    HIP_OK(hipMalloc(&dacc, 6 * 64 * sizeof(double)));
    fr.p.push_back(dacc);
    HIP_OK(hipMalloc(&dmax, 2 * 64 * sizeof(unsigned long long)));
    fr.p.push_back(dmax);
    dcnt = dmax + 64;
    HIP_OK(hipMemcpyAsync(dpi, pinfo.data(), pinfo.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    k_fill<<<grid_for(6 * 64, 256), 256, 0, st>>>(dacc, 6 * 64, 0.0);
    k_fill<<<1, 128, 0, st>>>(reinterpret_cast<double*>(dmax), 128, 0.0);
    HIP_OK(hipGetLastError());
    TaskSet& ts = get_tasks(pl, thr);
    f.pinfo = dpi;
    f.acc = dacc;
    f.maxbits = dmax;
    f.count = dcnt;
    for (int X = 5; X >= 0; --X)
        for (int Y = X; Y >= 0; --Y) {
            if (ts.cnt[X][Y] == 0) continue;
            const int2* tq = ensure_team_quartets(pl, ts, X, Y);
            const long long nq = ts.cum_sq[X][Y].empty() ? 0 : ts.cum_sq[X][Y].back();
            if (!tq || nq <= 0) continue;
            f.LA = kClassL[X][0];
            f.LB = kClassL[X][1];
            f.LC = kClassL[Y][0];
            f.LD = kClassL[Y][1];
            const long long nblk = static_cast<long long>(ncart(f.LA)) * ncart(f.LB) * ncart(f.LC) * ncart(f.LD);
            const long long batch = std::min(nq, std::max(1ll, kBatchDoubles / nblk));
            if (!dblk) {
                HIP_OK(hipMalloc(&dblk, kBatchDoubles * sizeof(double)));
                fr.p.push_back(dblk);
            }
            for (long long b0 = 0; b0 < nq; b0 += batch) {
                const int nb = static_cast<int>(std::min(batch, nq - b0));
                StoreArgs sa{};
                sa.pd = pl->pair_data();
                sa.tb = ctx.tb;
                sa.quartets = tq + b0;
                sa.off = nullptr;
                sa.stride = nblk;
                sa.n = nb;
                sa.out = dblk;
                launch_store(g_kernels[X][Y], sa, st);
                f.blocks = dblk;
                f.quartets = tq + b0;
                f.n = nb;
                k_fingerprint<<<static_cast<unsigned>(nb), 64, 0, st>>>(f);
                HIP_OK(hipGetLastError());
            }
        }
    std::vector<double> hacc(6 * 64);
    std::vector<unsigned long long> hm(128);
    HIP_OK(hipMemcpyAsync(hacc.data(), dacc, hacc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hm.data(), dmax, hm.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    long double tot[6] = {};
    double mx = 0.0;
    unsigned long long c = 0;
    for (int k = 0; k < 64; ++k) {
        for (int j = 0; j < 6; ++j) tot[j] += hacc[6 * k + j];
        double m;
        std::memcpy(&m, &hm[k], sizeof(double));
        mx = std::max(mx, m);
        c += hm[64 + k];
    }
    for (int j = 0; j < 6; ++j) out[j] = static_cast<double>(tot[j]);
    out[6] = mx;
    *count = c;
}

// Per class pair, time the lane and the team evaluator on this plan (FP64,
// J and K, a constant density, serialized launches bracketed by events) and
// keep the faster one; returns a JSON summary.
// Restores the plan settings the autotuner changes, and frees its scratch
// buffers, however the autotuner ends: on an error the plan keeps its former
// kernel choices and screening settings.
struct AutotuneScope {
    tlno_gpu_plan* pl;
    int timing, learned;
    double gate, dthr, fastschwarz, decay;
    signed char team_sel[6][6];
    bool fp32_slower[6][6];
    double *dD = nullptr, *dJ = nullptr, *dK = nullptr;
    bool done = false;
    explicit AutotuneScope(tlno_gpu_plan* p)
        : pl(p), timing(p->timing), learned(p->learned), gate(p->fp32_gate), dthr(p->density_thr),
          fastschwarz(p->fastschwarz), decay(p->decay_thr) {
        std::memcpy(team_sel, p->team_sel, sizeof(team_sel));
        std::memcpy(fp32_slower, p->fp32_slower, sizeof(fp32_slower));
    }
    ~AutotuneScope() {
        if (!done) {  // error path: the former choices stand
            std::memcpy(pl->team_sel, team_sel, sizeof(team_sel));
            std::memcpy(pl->fp32_slower, fp32_slower, sizeof(fp32_slower));
            (void)hipStreamSynchronize(pl->stream);
        }
        pl->timing = timing;
        std::memset(pl->cls_ms, 0, sizeof(pl->cls_ms));
        pl->timed_builds = 0;
        pl->fp32_gate = gate;
        pl->density_thr = dthr;
        pl->learned = learned;
        pl->fastschwarz = fastschwarz;
        pl->decay_thr = decay;
        (void)hipFree(dD);
        (void)hipFree(dJ);
        (void)hipFree(dK);
    }
};

// A lane class's FP32 share runs the FP32 kernels only where they beat the
// FP64 kernels on the whole class by this fraction: routed builds then split
// the class into two task lists, which fill their tasks worse than one (on
// water40's screened build, classes whose FP32 kernels were 3-5 % faster ran
// 3-6 % slower routed than kept in FP64).
#ifndef TLNO_FP32_MIN_GAIN
#define TLNO_FP32_MIN_GAIN 0.10
#endif
constexpr double kFp32MinGain = TLNO_FP32_MIN_GAIN;

std::string plan_autotune(tlno_gpu_plan* pl, double thr, int repeats) {
    HIP_OK(hipSetDevice(pl->device));
    const long mo2 = static_cast<long>(pl->Mout) * pl->Mout;
    AutotuneScope scope(pl);
    HIP_OK(hipMalloc(&scope.dD, mo2 * sizeof(double)));
    HIP_OK(hipMalloc(&scope.dJ, mo2 * sizeof(double)));
    HIP_OK(hipMalloc(&scope.dK, mo2 * sizeof(double)));
    double *dD = scope.dD, *dJ = scope.dJ, *dK = scope.dK;
    k_fill<<<grid_for(mo2, 256), 256, 0, pl->stream>>>(dD, mo2, 1e-3);
    pl->fp32_gate = 0.0;
    pl->density_thr = 0.0;
    pl->learned = 0;
    pl->fastschwarz = 0.0;
    pl->decay_thr = 0.0;
    if (!pl->timing)
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y < 6; ++Y)
                for (int e = 0; e < 2; ++e)
                    if (!pl->ev[X][Y][e]) HIP_OK(hipEventCreate(&pl->ev[X][Y][e]));
    pl->timing = 1;
    double ms[3][6][6] = {};
    std::memset(pl->fp32_slower, 0, sizeof(pl->fp32_slower));
    // modes: 0 lane (FP64), 1 team, 2 lane with every quartet on the FP32 kernels
    for (int mode = 0; mode < 3; ++mode) {
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y < 6; ++Y) pl->team_sel[X][Y] = static_cast<signed char>(mode == 1 ? 1 : 0);
        pl->fp32_gate = mode == 2 ? 1e300 : 0.0;
        jk_device(pl, dD, dJ, dK, thr, plan_buffers(pl));  // warm-up
        std::memset(pl->cls_ms, 0, sizeof(pl->cls_ms));
        pl->timed_builds = 0;
        for (int r = 0; r < repeats; ++r) jk_device(pl, dD, dJ, dK, thr, plan_buffers(pl));
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y < 6; ++Y) ms[mode][X][Y] = pl->cls_ms[X][Y] / std::max<long long>(1, pl->timed_builds);
        // test hook: an error in the middle of the autotuner (after the team
        // mode changed every class choice), for the restore guarantee
        if (mode == 1 && std::getenv("TLNO_TEST_AUTOTUNE_FAIL"))
            throw std::runtime_error("tlno_gpu: autotune interrupted (TLNO_TEST_AUTOTUNE_FAIL)");
    }
    std::string rep = "{";
    bool first = true;
    static const char* names[6] = {"ss", "ps", "pp", "ds", "dp", "dd"};
    for (int X = 5; X >= 0; --X)
        for (int Y = X; Y >= 0; --Y) {
            const bool has_team = g_kernels[X][Y].team != nullptr;
            const bool team = has_team && ms[1][X][Y] < ms[0][X][Y];
            pl->team_sel[X][Y] = static_cast<signed char>(team ? 1 : 0);
            pl->cls_cost[X][Y] = team ? ms[1][X][Y] : ms[0][X][Y];
            pl->fp32_slower[X][Y] = ms[2][X][Y] > 0.0 && ms[2][X][Y] >= (1.0 - kFp32MinGain) * ms[0][X][Y];
            if (ms[0][X][Y] == 0.0) continue;
            char buf[256];
            if (has_team)
                std::snprintf(buf, sizeof(buf),
                              "%s\"(%s|%s)\": {\"lane_ms\": %.4f, \"team_ms\": %.4f, \"lane_fp32_ms\": %.4f, "
                              "\"choice\": \"%s\"}",
                              first ? "" : ", ", names[X], names[Y], ms[0][X][Y], ms[1][X][Y], ms[2][X][Y],
                              team ? "team" : "lane");
            else
                std::snprintf(buf, sizeof(buf),
                              "%s\"(%s|%s)\": {\"lane_ms\": %.4f, \"lane_fp32_ms\": %.4f, \"choice\": \"lane\"}",
                              first ? "" : ", ", names[X], names[Y], ms[0][X][Y], ms[2][X][Y]);
            rep += buf;
            first = false;
        }
    rep += "}";
    HIP_OK(hipStreamSynchronize(pl->stream));
    scope.done = true;
    return rep;
}

void validate_density(const double* D, int M) {
    const double tol = 128.0 * 2.220446049250313e-16;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j <= i; ++j) {
            const double a = D[static_cast<std::size_t>(i) * M + j], b = D[static_cast<std::size_t>(j) * M + i];
            if (!std::isfinite(a) || !std::isfinite(b))
                throw std::runtime_error("eri_ao_tensor_hgp: density contains a non-finite value.");
            const double sc = std::max({1.0, std::abs(a), std::abs(b)});
            if (std::abs(a - b) > tol * sc) throw std::runtime_error("eri_ao_tensor_hgp: density must be symmetric.");
        }
}


}  // namespace tlno_amd

namespace tlno_amd {

// One-electron group-pair screening of the reference
// (integrals/one_electron/onee_matrices.cpp:100-210): with c0 the stored
// coefficients of a group's first component, K_g = sum_i |c0_i| (pi/a_i)^{3/4}
// and a_min its smallest exponent, a pair is dropped when
// K_a K_b exp(-mu_min AB^2) < eps (mu_min from the two smallest exponents);
// a kept pair evaluates V only when its overlap bound
// sum_ij |c0a_i c0b_j| (pi/p)^{3/2} exp(-chi AB^2) reaches eps.  eps = 0
// keeps every pair with V.  The list is grouped by pair class (plan order).
void onee_select_pairs(tlno_gpu_plan* pl, double eps) {
    const int ng = static_cast<int>(pl->groups.size());
    std::vector<double> K(ng, 0.0), amin(ng, 0.0);
    for (int g = 0; g < ng; ++g) {
        const GroupH& G = pl->groups[g];
        double a0 = G.exps[0], k = 0.0;
        for (int i = 0; i < G.nprim; ++i) {
            a0 = std::min(a0, G.exps[i]);
            k += std::abs(G.cr[i] * G.na_first) * std::pow(3.14159265358979323846 / G.exps[i], 0.75);
        }
        K[g] = k;
        amin[g] = a0;
    }
    std::vector<int2> list;
    list.reserve(pl->npair);
    long long kept = 0;
    for (int c = 0; c < kNumClasses; ++c) {
        pl->onee_cb[c] = static_cast<int>(list.size());
        for (int P = pl->cls_begin[c]; P < pl->cls_begin[c + 1]; ++P) {
            const GroupH& A = pl->groups[pl->pgA[P]];
            const GroupH& B = pl->groups[pl->pgB[P]];
            int do_v = 1;
            if (eps > 0.0) {
                double AB2 = 0.0;
                for (int x = 0; x < 3; ++x) AB2 += (A.C[x] - B.C[x]) * (A.C[x] - B.C[x]);
                const int ga = pl->pgA[P], gb = pl->pgB[P];
                const double mu = amin[ga] * amin[gb] / (amin[ga] + amin[gb]);
                if (K[ga] * K[gb] * std::exp(-mu * AB2) < eps) continue;
                double bound = 0.0;
                for (int i = 0; i < A.nprim; ++i)
                    for (int j = 0; j < B.nprim; ++j) {
                        const double p = A.exps[i] + B.exps[j];
                        const double chi = A.exps[i] * B.exps[j] / p;
                        bound += std::abs(A.cr[i] * A.na_first * B.cr[j] * B.na_first) * std::pow(3.14159265358979323846 / p, 1.5) *
                                 std::exp(-chi * AB2);
                    }
                do_v = bound >= eps;
            }
            list.push_back(make_int2(P, do_v));
            ++kept;
        }
    }
    pl->onee_cb[kNumClasses] = static_cast<int>(list.size());
    if (pl->d_onee_list) HIP_OK(hipFree(pl->d_onee_list));
    pl->d_onee_list = nullptr;
    HIP_OK(hipMalloc(&pl->d_onee_list, std::max<std::size_t>(1, list.size()) * sizeof(int2)));
    if (!list.empty())
        HIP_OK(hipMemcpy(pl->d_onee_list, list.data(), list.size() * sizeof(int2), hipMemcpyHostToDevice));
    pl->onee_list = std::move(list);
    pl->onee_kept = kept;
    pl->onee_decay = eps;
}

// Nuclear primitive-pair accounting of the reference's
// tlno_onee_nuclear_pair_stats (benchmarks/one_electron_engine/src/
// onee_capi.cpp:183-209): over kept pairs that evaluate V, all primitive
// pairs, and those whose monopole estimate (2 pi/p) exp(-chi AB^2)
// max|c_a| max|c_b| reaches thr.
void onee_nuclear_counts(const tlno_gpu_plan* pl, double thr, long long* total, long long* kept) {
    long long t = 0, k = 0;
    for (const int2& e : pl->onee_list) {
        if (!e.y) continue;
        const GroupH& A = pl->groups[pl->pgA[e.x]];
        const GroupH& B = pl->groups[pl->pgB[e.x]];
        const double na = onee_namax(A.L), nb = onee_namax(B.L);
        double AB2 = 0.0;
        for (int x = 0; x < 3; ++x) AB2 += (A.C[x] - B.C[x]) * (A.C[x] - B.C[x]);
        t += static_cast<long long>(A.nprim) * B.nprim;
        for (int i = 0; i < A.nprim; ++i)
            for (int j = 0; j < B.nprim; ++j) {
                const double p = A.exps[i] + B.exps[j];
                const double chi = A.exps[i] * B.exps[j] / p;
                const double scr = 2.0 * 3.14159265358979323846 / p * std::exp(-chi * AB2) * std::abs(A.cr[i]) * na * std::abs(B.cr[j]) * nb;
                if (scr >= thr) ++k;
            }
    }
    *total = t;
    *kept = k;
}

}  // namespace tlno_amd

tlno_gpu_plan::~tlno_gpu_plan() {
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    f(dA); f(dAB); f(dpw); f(dao); f(dpoff); f(dpn); f(ddiag); f(dpP); f(dpexp); f(dpL);
    f(d_int2basis); f(d_intL); f(d_intloc); f(d_intsb); f(d_sphL); f(d_sphloc); f(d_sphib);
    f(dDin); f(dDint); f(dJacc); f(dKacc); f(dJout); f(dKout); f(d_scalar);
    f(d_row_group); f(d_pgA); f(d_pgB); f(d_pcls); f(d_clsb); f(d_Q); f(d_gmax); f(d_clsmax); f(d_pairfac);
    f(d_cut); f(d_counter); f(d_gwork);
    f(d_pf0); f(d_pf1); f(d_pcen); f(d_grows); f(d_blk); f(d_bsum); f(d_frag); f(d_mvec); f(d_mnorm);
    f(d_qlist); f(d_t64); f(d_t32); f(d_rstats); f(d_cand);
    f(d_onee_list); f(d_onee_X); f(d_onee_O); f(d_onee_nuc); f(d_decay);
    for (auto& e : onee_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& kv : tasks) {
        f(kv.second.d_tasks);
        f(kv.second.d_bk);
        f(kv.second.d_ltasks);
        f(kv.second.d_rows);
        f(kv.second.d_rbase);
        f(kv.second.d_rtask);
        f(kv.second.d_rcnt);
        f(kv.second.d_hq);
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y < 6; ++Y) f(kv.second.d_tq[X][Y]);
    }
    if (timing)
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y < 6; ++Y)
                for (int e = 0; e < 2; ++e)
                    if (ev[X][Y][e]) (void)hipEventDestroy(ev[X][Y][e]);
    for (auto& row : hev)
        for (auto& pr : row)
            for (auto& e : pr)
                if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

using namespace tlno_amd;

extern "C" {

const char* tlno_gpu_last_error(void) { return g_err.c_str(); }

int tlno_gpu_init(int device) {
    return guarded([&] {
        int n = 0;
        HIP_OK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw std::runtime_error("tlno_gpu: no such HIP device");
        g_device = device;
        device_ctx(device);
    });
}

int tlno_gpu_plan_create(int n_comp, const double* centers, const int* ang, const int* nprim,
                         const double* exps, const double* coefs, int spherical, tlno_gpu_plan** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("tlno_gpu: null output handle");
        *out = nullptr;
        if (n_comp <= 0 || !centers || !ang || !nprim || !exps || !coefs)
            throw std::runtime_error("eri_ao_tensor_hgp: empty or null basis.");
        auto pl = std::make_unique<tlno_gpu_plan>();
        pl->device = g_device;
        pl->spherical = spherical ? 1 : 0;
        build_plan(pl.get(), n_comp, centers, ang, nprim, exps, coefs);
        *out = pl.release();
    });
}

void tlno_gpu_plan_destroy(tlno_gpu_plan* plan) { delete plan; }

int tlno_gpu_plan_info(const tlno_gpu_plan* pl, int* M, int* n_groups, int* n_pairs, long long* n_prim_pairs) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (M) *M = pl->Mout;
        if (n_groups) *n_groups = static_cast<int>(pl->groups.size());
        if (n_pairs) *n_pairs = pl->npair;
        if (n_prim_pairs) *n_prim_pairs = pl->nprim_pairs;
    });
}

int tlno_gpu_set_timing(tlno_gpu_plan* pl, int enable) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        HIP_OK(hipSetDevice(pl->device));
        if (enable && !pl->timing)
            for (int X = 0; X < 6; ++X)
                for (int Y = 0; Y < 6; ++Y)
                    for (int e = 0; e < 2; ++e) HIP_OK(hipEventCreate(&pl->ev[X][Y][e]));
        if (enable)
            for (int h = 0; h < kNumHigh; ++h)
                for (int Y = 0; Y < kNumClasses; ++Y)
                    for (int e = 0; e < 2; ++e)
                        if (!pl->hev[h][Y][e]) HIP_OK(hipEventCreate(&pl->hev[h][Y][e]));
        pl->timing = enable ? 1 : 0;
        std::memset(pl->cls_ms, 0, sizeof(pl->cls_ms));
        std::memset(pl->hcls_ms, 0, sizeof(pl->hcls_ms));
        pl->timed_builds = 0;
    });
}

int tlno_gpu_class_stats_all(const tlno_gpu_plan* pl, long long* shellq, long long* primq, double* ms,
                             long long* builds) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        auto it = pl->tasks.find(pl->last_thr);
        const bool have = it != pl->tasks.end();
        for (int X = 0; X < kNumClasses; ++X)
            for (int Y = 0; Y < kNumClasses; ++Y) {
                const int i = kNumClasses * X + Y;
                long long sq = 0, pq = 0;
                double m = 0.0;
                if (X < kNumLowClasses && Y < kNumLowClasses) {
                    if (have) {
                        sq = it->second.cls_sq[X][Y];
                        pq = it->second.cls_pq[X][Y];
                    }
                    m = pl->cls_ms[X][Y];
                } else if (X >= kNumLowClasses && Y <= X) {
                    if (have) {
                        sq = it->second.hcnt[X - kNumLowClasses][Y];
                        pq = it->second.hpq[X - kNumLowClasses][Y];
                    }
                    m = pl->hcls_ms[X - kNumLowClasses][Y];
                }
                if (shellq) shellq[i] = sq;
                if (primq) primq[i] = pq;
                if (ms) ms[i] = m;
            }
        if (builds) *builds = pl->timed_builds;
    });
}

int tlno_gpu_class_stats(const tlno_gpu_plan* pl, long long* shellq, long long* primq, double* ms,
                         long long* builds) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        auto it = pl->tasks.find(pl->last_thr);
        for (int X = 0; X < 6; ++X)
            for (int Y = 0; Y < 6; ++Y) {
                const int i = 6 * X + Y;
                if (shellq) shellq[i] = it != pl->tasks.end() ? it->second.cls_sq[X][Y] : 0;
                if (primq) primq[i] = it != pl->tasks.end() ? it->second.cls_pq[X][Y] : 0;
                if (ms) ms[i] = pl->cls_ms[X][Y];
            }
        if (builds) *builds = pl->timed_builds;
    });
}

int tlno_gpu_plan_pairs(const tlno_gpu_plan* pl, int* group_a, int* group_b, double* Q, int* class_begin) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        for (int k = 0; k < pl->npair; ++k) {
            if (group_a) group_a[k] = pl->pgA[k];
            if (group_b) group_b[k] = pl->pgB[k];
            if (Q) Q[k] = pl->Q[k];
        }
        if (class_begin)
            for (int c = 0; c <= kNumClasses; ++c) class_begin[c] = pl->cls_begin[c];
    });
}

int tlno_gpu_quartet_maxabs(tlno_gpu_plan* pl, int n, const int* pp, const int* pq, double* out) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        auto cls_of_pair = [&](int P) {
            int c = 0;
            while (!(P >= pl->cls_begin[c] && P < pl->cls_begin[c + 1])) ++c;
            return c;
        };
        auto nval = [&](int P, int Q) {
            return static_cast<long long>(ncart(pl->groups[pl->pgA[P]].L)) * ncart(pl->groups[pl->pgB[P]].L) *
                   ncart(pl->groups[pl->pgA[Q]].L) * ncart(pl->groups[pl->pgB[Q]].L);
        };
        std::map<std::pair<int, int>, std::vector<int>> by_class;
        for (int i = 0; i < n; ++i) {
            if (pp[i] < 0 || pp[i] >= pl->npair || pq[i] < 0 || pq[i] >= pl->npair)
                throw std::runtime_error("tlno_gpu: pair index out of range");
            int X = cls_of_pair(pp[i]), Y = cls_of_pair(pq[i]);
            by_class[{std::max(X, Y), std::min(X, Y)}].push_back(i);
        }
        for (auto& kv : by_class) {
            const auto& idx = kv.second;
            for (std::size_t b0 = 0; b0 < idx.size(); b0 += 262144) {
                const std::size_t b1 = std::min(idx.size(), b0 + 262144);
                std::vector<int2> qs;
                std::vector<long long> off;
                long long total = 0;
                for (std::size_t j = b0; j < b1; ++j) {
                    int P = pp[idx[j]], Q = pq[idx[j]];
                    if (cls_of_pair(P) < cls_of_pair(Q)) std::swap(P, Q);
                    qs.push_back(make_int2(P, Q));
                    off.push_back(total);
                    total += nval(P, Q);
                }
                off.push_back(total);
                double* dout = nullptr;
                HIP_OK(hipMalloc(&dout, static_cast<std::size_t>(total) * sizeof(double)));
                std::vector<long long> offq(off.begin(), off.end() - 1);
                run_store(pl, kv.first.first, kv.first.second, qs, offq, dout);
                std::vector<double> h(static_cast<std::size_t>(total));
                HIP_OK(hipMemcpy(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost));
                HIP_OK(hipFree(dout));
                for (std::size_t j = b0; j < b1; ++j) {
                    double m = 0.0;
                    for (long long t = off[j - b0]; t < off[j - b0 + 1]; ++t) m = std::max(m, std::abs(h[t]));
                    out[idx[j]] = m;
                }
            }
        }
    });
}

int tlno_gpu_set_density_threshold(tlno_gpu_plan* pl, double density_threshold) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!std::isfinite(density_threshold) || density_threshold < 0.0)
            throw std::runtime_error("eri_ao_tensor_hgp: density_threshold must be finite and >= 0.");
        pl->density_thr = density_threshold;
    });
}

int tlno_gpu_set_primitive_screens(tlno_gpu_plan* pl, double primitive_fastschwarz_threshold,
                                   double decay_threshold) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!std::isfinite(primitive_fastschwarz_threshold) || primitive_fastschwarz_threshold < 0.0)
            throw std::runtime_error("eri_ao_tensor_hgp: primitive_fastschwarz_threshold must be finite and >= 0.");
        if (!std::isfinite(decay_threshold) || decay_threshold < 0.0)
            throw std::runtime_error("eri_ao_tensor_hgp: decay_threshold must be finite and >= 0.");
        pl->fastschwarz = primitive_fastschwarz_threshold;
        pl->decay_thr = decay_threshold;
    });
}

int tlno_gpu_set_team_threshold(tlno_gpu_plan* pl, int min_outputs) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (min_outputs < 0) throw std::runtime_error("tlno_gpu: team threshold must be >= 0");
        pl->team_min_out = min_outputs;
        std::memset(pl->team_sel, -1, sizeof(pl->team_sel));  // an explicit policy replaces autotuned choices
        std::memset(pl->fp32_slower, 0, sizeof(pl->fp32_slower));
    });
}

int tlno_gpu_plan_autotune(tlno_gpu_plan* pl, double schwarz, int repeats, char* report, int report_len) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!std::isfinite(schwarz) || schwarz < 0.0)
            throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
        const std::string r = plan_autotune(pl, schwarz, std::max(1, repeats));
        if (report && report_len > 0) {
            std::strncpy(report, r.c_str(), static_cast<std::size_t>(report_len) - 1);
            report[report_len - 1] = '\0';
        }
    });
}

int tlno_gpu_set_mixed_precision(tlno_gpu_plan* pl, double fp32_threshold) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!std::isfinite(fp32_threshold) || fp32_threshold < 0.0)
            throw std::runtime_error("tlno_gpu: fp32 threshold must be finite and >= 0");
        pl->fp32_gate = fp32_threshold;
    });
}

int tlno_gpu_quartet_features(tlno_gpu_plan* pl, int n, const int* pair_p, const int* pair_q, float* out) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (n < 0 || (n > 0 && (!pair_p || !pair_q || !out))) throw std::runtime_error("tlno_gpu: bad feature arguments");
        for (int i = 0; i < n; ++i)
            if (pair_p[i] < 0 || pair_p[i] >= pl->npair || pair_q[i] < 0 || pair_q[i] >= pl->npair)
                throw std::runtime_error("tlno_gpu: pair index out of range");
        const int nlow = pl->cls_begin[kNumLowClasses];
        for (int i = 0; i < n; ++i)
            if (pair_p[i] >= nlow || pair_q[i] >= nlow)
                throw std::runtime_error("tlno_gpu: the screening network's features cover shells up to d");
        if (n == 0) return;
        HIP_OK(hipSetDevice(pl->device));
        int *dp = nullptr, *dq = nullptr;
        float* dout = nullptr;
        HIP_OK(hipMalloc(&dp, n * sizeof(int)));
        HIP_OK(hipMalloc(&dq, n * sizeof(int)));
        HIP_OK(hipMalloc(&dout, static_cast<std::size_t>(n) * MLP_IN * sizeof(float)));
        HIP_OK(hipMemcpy(dp, pair_p, n * sizeof(int), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dq, pair_q, n * sizeof(int), hipMemcpyHostToDevice));
        k_features<<<grid_for(n, 256), 256, 0, pl->stream>>>(FeatArgs{pl->d_pf0, pl->d_pf1, pl->d_pcen}, pl->d_pcls,
                                                             dp, dq, n, dout);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(pl->stream));
        HIP_OK(hipMemcpy(out, dout, static_cast<std::size_t>(n) * MLP_IN * sizeof(float), hipMemcpyDeviceToHost));
        HIP_OK(hipFree(dp));
        HIP_OK(hipFree(dq));
        HIP_OK(hipFree(dout));
    });
}

int tlno_gpu_eri_checksum(tlno_gpu_plan* pl, double schwarz, double* checksum) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("eri_ao_tensor_hgp: null packed ERI plan.");
        if (!checksum) throw std::runtime_error("tlno_gpu: null output");
        *checksum = eri_checksum(pl, schwarz);
    });
}

int tlno_gpu_eri_fingerprint(tlno_gpu_plan* pl, double schwarz, double* out, unsigned long long* count) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("eri_ao_tensor_hgp: null packed ERI plan.");
        if (!out || !count) throw std::runtime_error("tlno_gpu: null output");
        eri_fingerprint(pl, schwarz, out, count);
    });
}

int tlno_gpu_set_learned_screen(tlno_gpu_plan* pl, const float* params, double margin) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!params) {
            pl->learned = 0;
            return;
        }
        if (!std::isfinite(margin)) throw std::runtime_error("tlno_gpu: learned-screen margin must be finite");
        for (int i = 0; i < MLP_NALL; ++i)
            if (!std::isfinite(params[i])) throw std::runtime_error("tlno_gpu: non-finite network parameter");
        HIP_OK(hipSetDevice(pl->device));
        std::vector<uint4> frag(MLP_NFRAG * 64);
        std::vector<float> vec(192), norm(64);
        float b3 = 0.0f;
        mlp_pack(params, frag.data(), vec.data(), norm.data(), &b3);
        HIP_OK(hipStreamSynchronize(pl->stream));
        dev_upload(pl->d_frag, frag);
        dev_upload(pl->d_mvec, vec);
        dev_upload(pl->d_mnorm, norm);
        pl->mb3 = b3;
        pl->learned_margin = margin;
        pl->learned = 1;
    });
}

int tlno_gpu_last_route_stats(const tlno_gpu_plan* pl, long long* stats) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!stats) throw std::runtime_error("tlno_gpu: null output");
        for (int i = 0; i < 4; ++i) stats[i] = pl->last_route[i];
    });
}

int tlno_gpu_last_mixed_stats(const tlno_gpu_plan* pl, long long* fp32_shell_quartets,
                              long long* fp32_prim_quartets, double* max_abs_density) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (fp32_shell_quartets) *fp32_shell_quartets = pl->last_fp32_sq;
        if (fp32_prim_quartets) *fp32_prim_quartets = pl->last_fp32_pq;
        if (max_abs_density) *max_abs_density = pl->last_maxD;
    });
}

int tlno_gpu_plan_groups(const tlno_gpu_plan* pl, int max_groups, int* L, int* comps) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        const int ng = static_cast<int>(pl->groups.size());
        if (ng > max_groups) throw std::runtime_error("tlno_gpu: group buffer too small");
        for (int g = 0; g < ng; ++g) {
            if (L) L[g] = pl->groups[g].L;
            if (comps)
                for (int k = 0; k < 15; ++k) comps[15 * g + k] = k < ncart(pl->groups[g].L) ? pl->groups[g].comp[k] : -1;
        }
    });
}

int tlno_gpu_build_jk(tlno_gpu_plan* pl, const double* density, double* J, double* K, double schwarz) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!density) throw std::runtime_error("eri_ao_tensor_hgp: density must be spherical M x M row-major.");
        const int M = pl->Mout;
        validate_density(density, M);
        const std::size_t bytes = static_cast<std::size_t>(M) * M * sizeof(double);
        HIP_OK(hipSetDevice(pl->device));
        HIP_OK(hipMemcpyAsync(pl->dDin, density, bytes, hipMemcpyHostToDevice, pl->stream));
        jk_device(pl, pl->dDin, J ? pl->dJout : nullptr, K ? pl->dKout : nullptr, schwarz, plan_buffers(pl));
        if (J) HIP_OK(hipMemcpy(J, pl->dJout, bytes, hipMemcpyDeviceToHost));
        if (K) HIP_OK(hipMemcpy(K, pl->dKout, bytes, hipMemcpyDeviceToHost));
    });
}

int tlno_gpu_build_jk_device(tlno_gpu_plan* pl, const double* dD, double* dJ, double* dK, double schwarz) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!dD) throw std::runtime_error("eri_ao_tensor_hgp: null density.");
        jk_device(pl, dD, dJ, dK, schwarz, plan_buffers(pl));
    });
}

int tlno_gpu_build_jk_device_noalloc(tlno_gpu_plan* pl, const double* dD, double* dJ, double* dK, double schwarz) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (!dD) throw std::runtime_error("tlno_gpu_build_jk_device_noalloc: density must not be null");
        if (pl->Mcart > kNoAllocMaxM)
            throw std::runtime_error("tlno_gpu_build_jk_device_noalloc: " + std::to_string(pl->Mcart) +
                                     " Cartesian functions exceed the scratch limit of " +
                                     std::to_string(kNoAllocMaxM) + " (3 M^2 doubles of scratch per device)");
        DeviceCtx& ctx = device_ctx(pl->device);  // set up when the plan was created: no allocation here
        const std::size_t mc2 = static_cast<std::size_t>(pl->Mcart) * pl->Mcart;
        BuildBuffers bb{ctx.scratch, ctx.scratch + mc2, ctx.scratch + 2 * mc2,
                        reinterpret_cast<unsigned long long*>(ctx.scratch + 3 * mc2),
                        ctx.scratch + 3 * mc2 + 4 * kCounterSlots};
        struct Flag {
            Flag() { g_no_alloc = true; }
            ~Flag() { g_no_alloc = false; }
        } flag;
        jk_device(pl, dD, dJ, dK, schwarz, bb);
    });
}

int tlno_gpu_last_build_stats(const tlno_gpu_plan* pl, long long* sq, long long* pq) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (sq) *sq = pl->last_evaluated;
        if (pq) *pq = pl->last_evaluated_prim;
    });
}

int tlno_gpu_eri_cart_blocks(tlno_gpu_plan* pl, int n, const int* g4, const long long* offsets, double* out) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        const int ng = static_cast<int>(pl->groups.size());
        // pair index of each unordered group pair
        std::map<std::pair<int, int>, int> pidx;
        for (int k = 0; k < pl->npair; ++k) pidx[{pl->pgA[k], pl->pgB[k]}] = k;
        auto cls_of_pair = [&](int P) {
            int c = 0;
            while (!(P >= pl->cls_begin[c] && P < pl->cls_begin[c + 1])) ++c;
            return c;
        };
        struct Item {
            int i, P, Q;
            bool swap_bra, swap_ket, swap_pq;
        };
        std::vector<Item> items;
        for (int i = 0; i < n; ++i) {
            int g[4] = {g4[4 * i], g4[4 * i + 1], g4[4 * i + 2], g4[4 * i + 3]};
            for (int k = 0; k < 4; ++k)
                if (g[k] < 0 || g[k] >= ng) throw std::runtime_error("tlno_gpu: group index out of range");
            auto find = [&](int x, int y, bool& sw) {
                auto it = pidx.find({x, y});
                if (it != pidx.end()) {
                    sw = false;
                    return it->second;
                }
                sw = true;
                return pidx.at({y, x});
            };
            Item it;
            it.i = i;
            int P = find(g[0], g[1], it.swap_bra);
            int Q = find(g[2], g[3], it.swap_ket);
            it.swap_pq = cls_of_pair(P) < cls_of_pair(Q);
            if (it.swap_pq) std::swap(P, Q);
            it.P = P;
            it.Q = Q;
            items.push_back(it);
        }
        // evaluate per class pair
        std::map<std::pair<int, int>, std::vector<int>> by_class;
        for (int k = 0; k < static_cast<int>(items.size()); ++k)
            by_class[{cls_of_pair(items[k].P), cls_of_pair(items[k].Q)}].push_back(k);
        for (auto& kv : by_class) {
            const int X = kv.first.first, Y = kv.first.second;
            std::vector<int2> qs;
            std::vector<long long> off;
            long long total = 0;
            for (int k : kv.second) {
                qs.push_back(make_int2(items[k].P, items[k].Q));
                off.push_back(total);
                const int LA = pl->groups[pl->pgA[items[k].P]].L, LB = pl->groups[pl->pgB[items[k].P]].L;
                const int LC = pl->groups[pl->pgA[items[k].Q]].L, LD = pl->groups[pl->pgB[items[k].Q]].L;
                total += static_cast<long long>(ncart(LA)) * ncart(LB) * ncart(LC) * ncart(LD);
            }
            double* dout = nullptr;
            HIP_OK(hipMalloc(&dout, static_cast<std::size_t>(total) * sizeof(double)));
            run_store(pl, X, Y, qs, off, dout);
            std::vector<double> h(static_cast<std::size_t>(total));
            HIP_OK(hipMemcpy(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost));
            HIP_OK(hipFree(dout));
            for (std::size_t j = 0; j < kv.second.size(); ++j) {
                const Item& it = items[kv.second[j]];
                const GroupH* G[4] = {&pl->groups[pl->pgA[it.P]], &pl->groups[pl->pgB[it.P]],
                                      &pl->groups[pl->pgA[it.Q]], &pl->groups[pl->pgB[it.Q]]};
                const int n0 = ncart(G[0]->L), n1 = ncart(G[1]->L), n2 = ncart(G[2]->L), n3 = ncart(G[3]->L);
                const double* src = h.data() + off[j];
                // requested order of the four shells: map back through the swaps
                // stored block axes: (A_P, B_P, A_Q, B_Q)
                // requested axes r0..r3 correspond to stored axes:
                int map[4];
                {
                    int bra[2] = {0, 1}, ket[2] = {2, 3};
                    if (it.swap_pq) {
                        bra[0] = 2; bra[1] = 3; ket[0] = 0; ket[1] = 1;
                    }
                    map[0] = it.swap_bra ? bra[1] : bra[0];
                    map[1] = it.swap_bra ? bra[0] : bra[1];
                    map[2] = it.swap_ket ? ket[1] : ket[0];
                    map[3] = it.swap_ket ? ket[0] : ket[1];
                }
                const int dims[4] = {n0, n1, n2, n3};
                const int req_g[4] = {g4[4 * it.i], g4[4 * it.i + 1], g4[4 * it.i + 2], g4[4 * it.i + 3]};
                int rd[4];
                for (int k = 0; k < 4; ++k) rd[k] = dims[map[k]];
                // component order: basis-appearance order within each group
                std::vector<int> ord[4];
                for (int k = 0; k < 4; ++k) {
                    const GroupH& Gr = pl->groups[req_g[k]];
                    std::vector<std::pair<int, int>> v;
                    for (int c = 0; c < ncart(Gr.L); ++c) v.push_back({Gr.comp[c], c});
                    std::sort(v.begin(), v.end());
                    for (auto& x : v) ord[k].push_back(x.second);  // local index of k-th appearing component
                }
                double* dst = out + offsets[it.i];
                int e[4];
                for (e[0] = 0; e[0] < rd[0]; ++e[0])
                    for (e[1] = 0; e[1] < rd[1]; ++e[1])
                        for (e[2] = 0; e[2] < rd[2]; ++e[2])
                            for (e[3] = 0; e[3] < rd[3]; ++e[3]) {
                                int s[4];
                                for (int k = 0; k < 4; ++k) s[map[k]] = ord[k][e[k]];
                                dst[((static_cast<long long>(e[0]) * rd[1] + e[1]) * rd[2] + e[2]) * rd[3] + e[3]] =
                                    src[((static_cast<long long>(s[0]) * n1 + s[1]) * n2 + s[2]) * n3 + s[3]];
                            }
            }
        }
    });
}

int tlno_gpu_onee(tlno_gpu_plan* pl, int n_nuc, const double* Z, const double* R, double* S, double* T,
                  double* V) {
    return tlno_gpu_onee_screened(pl, n_nuc, Z, R, 0.0, 0.0, S, T, V, nullptr);
}

int tlno_gpu_onee_screened(tlno_gpu_plan* pl, int n_nuc, const double* Z, const double* R, double decay_eps,
                           double nuclear_eps, double* S, double* T, double* V, double* stats) {
    return guarded([&] {
        if (!pl) throw std::runtime_error("jk_direct: null J/K plan backend.");
        if (n_nuc < 0 || (n_nuc > 0 && (!Z || !R))) throw std::runtime_error("tlno_gpu: invalid nuclei");
        for (int c = 0; c < n_nuc; ++c)
            if (!std::isfinite(Z[c]) || !std::isfinite(R[3 * c]) || !std::isfinite(R[3 * c + 1]) ||
                !std::isfinite(R[3 * c + 2]))
                throw std::runtime_error("tlno_gpu: nuclear charges and centres must be finite");
        if (!std::isfinite(decay_eps) || decay_eps < 0.0 || !std::isfinite(nuclear_eps) || nuclear_eps < 0.0)
            throw std::runtime_error("onee: screening thresholds must be finite and >= 0.");
        DeviceCtx& ctx = device_ctx(pl->device);
        HIP_OK(hipSetDevice(pl->device));
        const hipStream_t st = pl->stream;
        const long mc2 = static_cast<long>(pl->Mcart) * pl->Mcart;
        const long mo2 = static_cast<long>(pl->Mout) * pl->Mout;
        if (pl->onee_decay != decay_eps) onee_select_pairs(pl, decay_eps);
        if (!pl->d_onee_X) {
            HIP_OK(hipMalloc(&pl->d_onee_X, 3 * mc2 * sizeof(double)));
            HIP_OK(hipMalloc(&pl->d_onee_O, 3 * mo2 * sizeof(double)));
            HIP_OK(hipEventCreate(&pl->onee_ev[0]));
            HIP_OK(hipEventCreate(&pl->onee_ev[1]));
        }
        if (n_nuc > pl->onee_nuc_cap) {
            if (pl->d_onee_nuc) HIP_OK(hipFree(pl->d_onee_nuc));
            pl->d_onee_nuc = nullptr;
            pl->onee_nuc_cap = 0;
            HIP_OK(hipMalloc(&pl->d_onee_nuc, 4 * static_cast<std::size_t>(n_nuc) * sizeof(double)));
            pl->onee_nuc_cap = n_nuc;
        }
        double* dR = pl->d_onee_nuc;
        double* dZ = pl->d_onee_nuc + 3 * static_cast<std::size_t>(pl->onee_nuc_cap);
        if (n_nuc > 0) {
            HIP_OK(hipMemcpyAsync(dR, R, 3 * n_nuc * sizeof(double), hipMemcpyHostToDevice, st));
            HIP_OK(hipMemcpyAsync(dZ, Z, n_nuc * sizeof(double), hipMemcpyHostToDevice, st));
        }
        double zsum = 0.0;
        for (int c = 0; c < n_nuc; ++c) zsum += std::abs(Z[c]);
        const double vthr = zsum > 0.0 ? nuclear_eps / zsum : 0.0;
        double* dX = pl->d_onee_X;
        double* dO = pl->d_onee_O;
        k_fill<<<grid_for(3 * mc2, 256), 256, 0, st>>>(dX, 3 * mc2, 0.0);
        HIP_OK(hipEventRecord(pl->onee_ev[0], st));
        launch_onee(pl->pair_data(), ctx.tb, pl->dpL, pl->dpexp, dR, dZ, n_nuc, pl->d_onee_list, pl->onee_cb, vthr,
                    pl->Mcart, dX, dX + mc2, dX + 2 * mc2, st);
        HIP_OK(hipEventRecord(pl->onee_ev[1], st));
        const C2SArgs c2s{ctx.c2s};
        for (int m = 0; m < 3; ++m) {
            if (pl->spherical)
                k_fin_sph<<<grid_for(mo2, 256), 256, 0, st>>>(dX + m * mc2, dO + m * mo2, pl->d_sphL, pl->d_sphloc,
                                                              pl->d_sphib, pl->Mcart, pl->Msph, c2s);
            else
                k_fin_cart<<<grid_for(mo2, 256), 256, 0, st>>>(dX + m * mc2, dO + m * mo2, pl->d_int2basis,
                                                               pl->Mcart);
        }
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));
        double* outs[3] = {S, T, V};
        for (int m = 0; m < 3; ++m)
            if (outs[m]) HIP_OK(hipMemcpy(outs[m], dO + m * mo2, mo2 * sizeof(double), hipMemcpyDeviceToHost));
        if (stats) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, pl->onee_ev[0], pl->onee_ev[1]));
            long long tot = 0, kept = 0;
            onee_nuclear_counts(pl, vthr, &tot, &kept);
            const double ng = static_cast<double>(pl->groups.size());
            stats[0] = ng;
            stats[1] = ng * (ng + 1.0) / 2.0;
            stats[2] = static_cast<double>(pl->onee_kept);
            stats[3] = static_cast<double>(tot);
            stats[4] = static_cast<double>(kept);
            stats[5] = ms;
        }
    });
}

int tlno_gpu_boys(int n, const double* T, int mmax, double* out) {
    return guarded([&] {
        if (n <= 0) return;
        if (mmax < 0 || mmax > TLNO_GPU_BOYS_MMAX)
            throw std::runtime_error("tlno_gpu: boys order must be 0.." + std::to_string(TLNO_GPU_BOYS_MMAX));
        if (!T || !out) throw std::runtime_error("tlno_gpu: boys arguments and output must not be null");
        for (int i = 0; i < n; ++i)
            if (!(std::isfinite(T[i]) && T[i] >= 0.0))
                throw std::runtime_error("tlno_gpu: boys argument must be finite and >= 0");
        DeviceCtx& ctx = device_ctx(g_device);
        HIP_OK(hipSetDevice(g_device));
        const std::size_t nout = static_cast<std::size_t>(n) * (mmax + 1);
        std::lock_guard<std::mutex> lk(ctx.util_mu);
        double* dT = ctx.util_buffer(n + nout);
        double* dO = dT + n;
        HIP_OK(hipMemcpy(dT, T, n * sizeof(double), hipMemcpyHostToDevice));
        const unsigned g = grid_for(n, 128);
        switch (mmax) {
#define TLNO_BOYS_CASE(m) \
    case m: k_boys<m><<<g, 128>>>(ctx.tb, dT, dO, n); break;
            TLNO_BOYS_CASE(0) TLNO_BOYS_CASE(1) TLNO_BOYS_CASE(2) TLNO_BOYS_CASE(3) TLNO_BOYS_CASE(4)
            TLNO_BOYS_CASE(5) TLNO_BOYS_CASE(6) TLNO_BOYS_CASE(7) TLNO_BOYS_CASE(8) TLNO_BOYS_CASE(9)
            TLNO_BOYS_CASE(10) TLNO_BOYS_CASE(11) TLNO_BOYS_CASE(12) TLNO_BOYS_CASE(13) TLNO_BOYS_CASE(14)
            TLNO_BOYS_CASE(15) TLNO_BOYS_CASE(16)
#undef TLNO_BOYS_CASE
            default: k_boys_any<<<g, 128>>>(dT, dO, n, mmax); break;
        }
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(out, dO, nout * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int tlno_gpu_rys(int N, int n, const double* T, double* roots, double* weights) {
    return guarded([&] {
        if (n <= 0) return;
        if (N < 1 || N > RYS_NMAX)
            throw std::runtime_error("tlno_gpu: Rys order must be 1.." + std::to_string(RYS_NMAX));
        if (!T || !roots || !weights) throw std::runtime_error("tlno_gpu: Rys arguments and outputs must not be null");
        for (int i = 0; i < n; ++i)
            if (!(std::isfinite(T[i]) && T[i] >= 0.0))
                throw std::runtime_error("tlno_gpu: Rys argument must be finite and >= 0");
        DeviceCtx& ctx = device_ctx(g_device);
        HIP_OK(hipSetDevice(g_device));
        const std::size_t nr = static_cast<std::size_t>(n) * N;
        std::lock_guard<std::mutex> lk(ctx.util_mu);
        double* dT = ctx.util_buffer(n + 2 * nr);
        double* dR = dT + n;
        double* dW = dR + nr;
        HIP_OK(hipMemcpy(dT, T, n * sizeof(double), hipMemcpyHostToDevice));
        const unsigned g = grid_for(n, 128);
        switch (N) {
            case 1: k_rys<1><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 2: k_rys<2><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 3: k_rys<3><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 4: k_rys<4><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 5: k_rys<5><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 6: k_rys<6><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 7: k_rys<7><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 8: k_rys<8><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
            case 9: k_rys<9><<<g, 128>>>(ctx.tb, dT, dR, dW, n); break;
        }
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(roots, dR, nr * sizeof(double), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(weights, dW, nr * sizeof(double), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

// ============================================================ molecule batches ===
// Many molecules share one set of class launches: their pair data are
// concatenated (class-major, then molecule, then descending Schwarz bound),
// every task pairs a bra with kets of its own molecule, and the D / J / K
// blocks of molecule m are packed at offset mbase[m] with row stride M_m
// (Cartesian component basis).  Class kernels run persistent (a fixed grid of
// waves striding over the class's tasks) with per-class launch shapes from
// the autotuner; a build can be captured once into a HIP graph and replayed.
namespace tlno_amd {

// element idx of the packed arrays -> molecule (mbase has nmol + 1 entries)
__device__ inline int batch_mol(const long long* mbase, int nmol, long long idx) {
    int lo = 0, hi = nmol;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (mbase[mid] <= idx) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_prep_batch(const double* __restrict__ Din, double* __restrict__ Dint, const int* __restrict__ perm,
                             const int* __restrict__ pbase, const long long* __restrict__ mbase,
                             const int* __restrict__ mM, int nmol, long long total) {
    const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int m = batch_mol(mbase, nmol, idx);
    const long long base = mbase[m];
    const int M = mM[m];
    const long long loc = idx - base;
    const int i = static_cast<int>(loc / M), j = static_cast<int>(loc % M);
    const int* pm = perm + pbase[m];
    Dint[idx] = Din[base + static_cast<long long>(pm[i]) * M + pm[j]];
}

__global__ void k_fin_batch(const double* __restrict__ Xacc, double* __restrict__ Xout, const int* __restrict__ perm,
                            const int* __restrict__ pbase, const long long* __restrict__ mbase,
                            const int* __restrict__ mM, int nmol, long long total) {
    const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int m = batch_mol(mbase, nmol, idx);
    const long long base = mbase[m];
    const int M = mM[m];
    const long long loc = idx - base;
    const int i = static_cast<int>(loc / M), j = static_cast<int>(loc % M);
    const int* pm = perm + pbase[m];
    Xout[base + static_cast<long long>(pm[i]) * M + pm[j]] = Xacc[idx] + Xacc[base + static_cast<long long>(j) * M + i];
}

// launch shape of one class: block size and persistent waves per CU (0: one
// wave per work item)
struct LaunchShape {
    int block = 256;
    int waves_per_cu = 0;
};

}  // namespace tlno_amd

struct tlno_gpu_batch {
    int device = 0;
    int nmol = 0;
    std::vector<std::unique_ptr<tlno_gpu_plan>> mols;
    std::vector<long long> mbase;  // nmol + 1
    std::vector<int> mM, pbase;
    long long total = 0;
    int npair = 0;
    int cls_begin[7] = {0};
    std::vector<int> pmol, pn;
    std::vector<double> Q;
    std::vector<int> mcls[7];  // mcls[c][m]: first pair of molecule m in class c (size nmol + 1)
    // device data
    double *dA = nullptr, *dAB = nullptr, *dpw = nullptr;
    int *dao = nullptr, *dpoff = nullptr, *dpn = nullptr, *ddiag = nullptr, *dpmol = nullptr;
    double4* dpP = nullptr;
    long long* dmbase = nullptr;
    int *dmM = nullptr, *dperm = nullptr, *dpbase = nullptr;
    double *dDint = nullptr, *dJacc = nullptr, *dKacc = nullptr;
    unsigned long long* dcounter = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int n_cu = 256;
    // work lists per Schwarz threshold
    struct Tasks {
        int4* d = nullptr;
        long long off[6][6] = {}, cnt[6][6] = {};
        long long shell_quartets = 0, prim_quartets = 0;
    };
    std::map<double, Tasks> tasks;
    tlno_amd::LaunchShape shape[6][6];
    double cost[6][6] = {};  // autotuned ms per class pair (0: not measured)
    int team_min_out = TLNO_TEAM_MIN_OUT;
    // captured build
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void* graph_key[3] = {nullptr, nullptr, nullptr};
    double graph_thr = -1.0;
    long long last_sq = 0, last_pq = 0;
    double last_ms = 0.0;

    tlno_amd::PairData pair_data() const {
        tlno_amd::PairData p;
        p.A = dA;
        p.AB = dAB;
        p.ao = dao;
        p.poff = dpoff;
        p.pn = dpn;
        p.pP = dpP;
        p.pw = dpw;
        p.diag = ddiag;
        return p;
    }
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_key[0] = graph_key[1] = graph_key[2] = nullptr;
        graph_thr = -1.0;
    }
    ~tlno_gpu_batch() {
        drop_graph();
        for (void* p : {(void*)dA, (void*)dAB, (void*)dpw, (void*)dao, (void*)dpoff, (void*)dpn, (void*)ddiag,
                        (void*)dpmol, (void*)dpP, (void*)dmbase, (void*)dmM, (void*)dperm, (void*)dpbase,
                        (void*)dDint, (void*)dJacc, (void*)dKacc, (void*)dcounter})
            if (p) (void)hipFree(p);
        for (auto& kv : tasks)
            if (kv.second.d) (void)hipFree(kv.second.d);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace tlno_amd {

void batch_build_tables(tlno_gpu_batch* bt) {
    const int nm = bt->nmol;
    // class-major, then molecule: pair order of the batch
    std::vector<double> A, AB, pw;
    std::vector<int> ao, poff, pn, diag, pmol, perm;
    std::vector<double4> pP;
    for (int c = 0; c < 6; ++c) {
        bt->cls_begin[c] = static_cast<int>(pn.size());
        bt->mcls[c].assign(nm + 1, 0);
        for (int m = 0; m < nm; ++m) {
            const tlno_gpu_plan& pl = *bt->mols[m];
            bt->mcls[c][m] = static_cast<int>(pn.size());
            for (int P = pl.cls_begin[c]; P < pl.cls_begin[c + 1]; ++P) {
                for (int x = 0; x < 3; ++x) {
                    A.push_back(pl.h_A[3 * P + x]);
                    AB.push_back(pl.h_AB[3 * P + x]);
                }
                ao.push_back(pl.h_ao[2 * P]);
                ao.push_back(pl.h_ao[2 * P + 1]);
                poff.push_back(static_cast<int>(pP.size()));
                pn.push_back(pl.pn_host[P]);
                diag.push_back(pl.h_diag[P]);
                pmol.push_back(m);
                bt->Q.push_back(pl.Q[P]);
                for (int k = 0; k < pl.pn_host[P]; ++k) {
                    pP.push_back(pl.h_pP[pl.h_poff[P] + k]);
                    pw.push_back(pl.h_pw[pl.h_poff[P] + k]);
                }
            }
        }
        bt->mcls[c][nm] = static_cast<int>(pn.size());
    }
    bt->cls_begin[6] = static_cast<int>(pn.size());
    bt->npair = static_cast<int>(pn.size());
    bt->pmol = pmol;
    bt->pn = pn;
    bt->mbase.assign(nm + 1, 0);
    bt->mM.assign(nm, 0);
    bt->pbase.assign(nm, 0);
    for (int m = 0; m < nm; ++m) {
        const tlno_gpu_plan& pl = *bt->mols[m];
        bt->mM[m] = pl.Mcart;
        bt->mbase[m + 1] = bt->mbase[m] + static_cast<long long>(pl.Mcart) * pl.Mcart;
        bt->pbase[m] = static_cast<int>(perm.size());
        perm.insert(perm.end(), pl.int2basis.begin(), pl.int2basis.end());
    }
    bt->total = bt->mbase[nm];
    dev_upload(bt->dA, A);
    dev_upload(bt->dAB, AB);
    dev_upload(bt->dpw, pw);
    dev_upload(bt->dao, ao);
    dev_upload(bt->dpoff, poff);
    dev_upload(bt->dpn, pn);
    dev_upload(bt->ddiag, diag);
    dev_upload(bt->dpmol, pmol);
    dev_upload(bt->dpP, pP);
    dev_upload(bt->dmbase, bt->mbase);
    dev_upload(bt->dmM, bt->mM);
    dev_upload(bt->dperm, perm);
    dev_upload(bt->dpbase, bt->pbase);
    const std::size_t tot = static_cast<std::size_t>(std::max(1ll, bt->total));
    HIP_OK(hipMalloc(&bt->dDint, tot * sizeof(double)));
    HIP_OK(hipMalloc(&bt->dJacc, tot * sizeof(double)));
    HIP_OK(hipMalloc(&bt->dKacc, tot * sizeof(double)));
    HIP_OK(hipMalloc(&bt->dcounter, 4 * kCounterSlots * sizeof(unsigned long long)));
}

tlno_gpu_batch::Tasks& batch_tasks(tlno_gpu_batch* bt, double thr) {
    auto it = bt->tasks.find(thr);
    if (it != bt->tasks.end()) return it->second;
    tlno_gpu_batch::Tasks ts;
    std::vector<int4> all;
    std::vector<long long> pref(bt->npair + 1, 0);
    for (int i = 0; i < bt->npair; ++i) pref[i + 1] = pref[i] + bt->pn[i];
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            ts.off[X][Y] = static_cast<long long>(all.size());
            std::vector<int4> base;
            for (int m = 0; m < bt->nmol; ++m) {
                const int k0 = bt->mcls[Y][m], k1 = bt->mcls[Y][m + 1];
                if (k1 <= k0) continue;
                for (int P = bt->mcls[X][m]; P < bt->mcls[X][m + 1]; ++P) {
                    int end = k1;
                    if (thr > 0.0) {
                        int lo = k0, hi = k1;
                        while (lo < hi) {
                            const int mid = (lo + hi) / 2;
                            if (bt->Q[P] * bt->Q[mid] >= thr) lo = mid + 1;
                            else hi = mid;
                        }
                        end = lo;
                    }
                    if (X == Y) end = std::min(end, P + 1);
                    for (int q0 = k0; q0 < end; q0 += 64) base.push_back(make_int4(P, q0, end, 0));
                    if (end > k0) {
                        ts.shell_quartets += end - k0;
                        ts.prim_quartets += (pref[end] - pref[k0]) * bt->pn[P];
                    }
                }
            }
            const long long waves = static_cast<long long>(base.size()) * g_kernels[X][Y].nchunk;
            constexpr long long kTargetWaves = TLNO_TARGET_WAVES;
            int split = 1;
            if (waves > 0 && waves < kTargetWaves)
                split = static_cast<int>(std::min<long long>(256, (kTargetWaves + waves - 1) / waves));
            for (const int4& t : base) {
                const int pbn = bt->pn[t.x];
                const int s = std::max(1, std::min(pbn, wave_split(split, std::min(t.z - t.y, 64))));
                const int step = (pbn + s - 1) / s;
                for (int i0 = 0; i0 < pbn; i0 += step)
                    all.push_back(make_int4(t.x, t.y, t.z, i0 | (std::min(pbn, i0 + step) << 16)));
            }
            ts.cnt[X][Y] = static_cast<long long>(all.size()) - ts.off[X][Y];
        }
    if (!all.empty()) {
        HIP_OK(hipMalloc(&ts.d, all.size() * sizeof(int4)));
        HIP_OK(hipMemcpy(ts.d, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    return bt->tasks.emplace(thr, ts).first->second;
}

JKArgs batch_args(tlno_gpu_batch* bt, bool do_j, bool do_k) {
    DeviceCtx& ctx = device_ctx(bt->device);
    JKArgs a{};
    a.pd = bt->pair_data();
    a.tb = ctx.tb;
    a.D = bt->dDint;
    a.J = bt->dJacc;
    a.K = bt->dKacc;
    a.M = 0;
    a.do_j = do_j;
    a.do_k = do_k;
    a.split = 1;
    a.counter = bt->dcounter;
    a.pmol = bt->dpmol;
    a.mbase = bt->dmbase;
    a.mM = bt->dmM;
    return a;
}

// one class's kernels with a given launch shape (range-mode tasks)
void batch_launch_class(tlno_gpu_batch* bt, const tlno_gpu_batch::Tasks& ts, int X, int Y, const LaunchShape& sh,
                        JKArgs a, hipStream_t st) {
    const long long nt = ts.cnt[X][Y];
    if (nt == 0) return;
    a.tasks = ts.d + ts.off[X][Y];
    a.ntasks = static_cast<int>(nt);
    long long waves = nt;
    if (sh.waves_per_cu > 0) {
        const long long grid_waves = std::min<long long>(nt, static_cast<long long>(bt->n_cu) * sh.waves_per_cu);
        a.stride = static_cast<int>(grid_waves);
        a.nwork = static_cast<int>(nt);
        waves = grid_waves;
    } else {
        a.stride = 0;
        a.nwork = 0;
    }
    const ClassKernels& k = g_kernels[X][Y];
    void* params[] = {&a};
    if (use_team(k, false, a, bt->team_min_out)) {  // one block per shell quartet, or a persistent grid
        const long long items = nt * 64;
        long long blocks = items;
        if (sh.waves_per_cu > 0) {
            blocks = std::min<long long>(items, static_cast<long long>(bt->n_cu) * sh.waves_per_cu);
            a.stride = static_cast<int>(blocks);
        } else {
            a.stride = 0;
        }
        HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(k.team), dim3(static_cast<unsigned>(blocks)), dim3(64),
                               params, 0, st));
        return;
    }
    const int block = sh.block;
    const unsigned grid = grid_for(waves * 64, block);
    for (int g = 0; g < k.ngroup; ++g)
        HIP_OK(hipLaunchKernel(reinterpret_cast<const void*>(k.grp[g].jk), dim3(grid, k.grp[g].n), dim3(block), params,
                               0, st));
}

// The whole build: prep and finalize on the batch stream, the class pairs
// list-scheduled onto the device's side streams between a fork and a join
// (capturable: kernels and event records/waits only).  The caller holds
// ctx.side_mu and has created the side streams.
void batch_enqueue(tlno_gpu_batch* bt, DeviceCtx& ctx, const double* dD, double* dJ, double* dK, double thr) {
    const tlno_gpu_batch::Tasks& ts = batch_tasks(bt, thr);
    const hipStream_t st = bt->stream;
    const unsigned g = grid_for(bt->total, 256);
    k_prep_batch<<<g, 256, 0, st>>>(dD, bt->dDint, bt->dperm, bt->dpbase, bt->dmbase, bt->dmM, bt->nmol, bt->total);
    if (dJ) k_fill<<<g, 256, 0, st>>>(bt->dJacc, bt->total, 0.0);
    if (dK) k_fill<<<g, 256, 0, st>>>(bt->dKacc, bt->total, 0.0);
    k_fill<<<grid_for(4 * kCounterSlots, 256), 256, 0, st>>>(reinterpret_cast<double*>(bt->dcounter), 4 * kCounterSlots, 0.0);
    const JKArgs a = batch_args(bt, dJ != nullptr, dK != nullptr);
    HIP_OK(hipEventRecord(ctx.ev_fork, st));
    for (int s = 0; s < ctx.nside; ++s) HIP_OK(hipStreamWaitEvent(ctx.side[s], ctx.ev_fork, 0));
    double load[kSideStreams] = {};
    for (int X = 5; X >= 0; --X)
        for (int Y = X; Y >= 0; --Y) {
            if (ts.cnt[X][Y] == 0) continue;
            // the autotuned time, else tasks x outputs per quartet
            const double c = bt->cost[X][Y] > 0.0
                                 ? bt->cost[X][Y]
                                 : 1e-6 * static_cast<double>(ts.cnt[X][Y]) * kClassNcart[X] * kClassNcart[Y];
            int best = 0;
            for (int q = 1; q < ctx.nside; ++q)
                if (load[q] < load[best]) best = q;
            load[best] += c;
            batch_launch_class(bt, ts, X, Y, bt->shape[X][Y], a, ctx.side[best]);
        }
    for (int s = 0; s < ctx.nside; ++s) {
        HIP_OK(hipEventRecord(ctx.ev_join[s], ctx.side[s]));
        HIP_OK(hipStreamWaitEvent(st, ctx.ev_join[s], 0));
    }
    if (dJ) k_fin_batch<<<g, 256, 0, st>>>(bt->dJacc, dJ, bt->dperm, bt->dpbase, bt->dmbase, bt->dmM, bt->nmol, bt->total);
    if (dK) k_fin_batch<<<g, 256, 0, st>>>(bt->dKacc, dK, bt->dperm, bt->dpbase, bt->dmbase, bt->dmM, bt->nmol, bt->total);
    HIP_OK(hipGetLastError());
}

void batch_build(tlno_gpu_batch* bt, const double* dD, double* dJ, double* dK, double thr, bool use_graph) {
    if (!std::isfinite(thr) || thr < 0.0)
        throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
    HIP_OK(hipSetDevice(bt->device));
    const tlno_gpu_batch::Tasks& ts = batch_tasks(bt, thr);  // host work stays outside any capture
    DeviceCtx& ctx = device_ctx(bt->device);
    // the side streams belong to this build until its capture ends or its
    // launches are enqueued
    std::lock_guard<std::mutex> side_lk(ctx.side_mu);
    ensure_side_streams(ctx);
    HIP_OK(hipEventRecord(bt->ev0, bt->stream));
    if (use_graph) {
        const bool same = bt->graph_exec && bt->graph_key[0] == dD && bt->graph_key[1] == dJ &&
                          bt->graph_key[2] == dK && bt->graph_thr == thr;
        if (!same) {
            bt->drop_graph();
            HIP_OK(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
            try {
                batch_enqueue(bt, ctx, dD, dJ, dK, thr);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(bt->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_OK(hipStreamEndCapture(bt->stream, &bt->graph));
            HIP_OK(hipGraphInstantiate(&bt->graph_exec, bt->graph, nullptr, nullptr, 0));
            bt->graph_key[0] = dD;
            bt->graph_key[1] = dJ;
            bt->graph_key[2] = dK;
            bt->graph_thr = thr;
        }
        HIP_OK(hipGraphLaunch(bt->graph_exec, bt->stream));
    } else {
        batch_enqueue(bt, ctx, dD, dJ, dK, thr);
    }
    HIP_OK(hipEventRecord(bt->ev1, bt->stream));
    HIP_OK(hipStreamSynchronize(bt->stream));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, bt->ev0, bt->ev1));
    bt->last_ms = ms;
    unsigned long long c[4] = {0, 0, 0, 0};
    std::vector<unsigned long long> slots(4 * kCounterSlots);
    HIP_OK(hipMemcpy(slots.data(), bt->dcounter, slots.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4 * kCounterSlots; ++k) c[k % 4] += slots[k];
    bt->last_sq = static_cast<long long>(c[0]);
    bt->last_pq = static_cast<long long>(c[1]);
    (void)ts;
}

// Per class: time every candidate launch shape on the class's own kernels
// (J and K digestion on scratch accumulators, the densities left as they
// are) and keep the fastest.  Candidates: block 64 / 128 / 256 threads x one
// wave per work item or 4 / 8 / 16 / 32 persistent waves per CU.
std::string batch_autotune(tlno_gpu_batch* bt, double thr, int repeats) {
    HIP_OK(hipSetDevice(bt->device));
    const tlno_gpu_batch::Tasks& ts = batch_tasks(bt, thr);
    const hipStream_t st = bt->stream;
    const unsigned g = grid_for(bt->total, 256);
    k_fill<<<g, 256, 0, st>>>(bt->dDint, bt->total, 1e-3);  // a smooth density for timing
    const JKArgs a = batch_args(bt, true, true);
    std::string rep = "{";
    bool first = true;
    for (int X = 5; X >= 0; --X)
        for (int Y = X; Y >= 0; --Y) {
            if (ts.cnt[X][Y] == 0) continue;
            LaunchShape best;
            double best_ms = 1e300, default_ms = 0.0;
            for (int block : {64, 128, 256})
                for (int wpc : {0, 4, 8, 16, 32}) {
                    const LaunchShape sh{block, wpc};
                    batch_launch_class(bt, ts, X, Y, sh, a, st);  // warm-up
                    HIP_OK(hipEventRecord(bt->ev0, st));
                    for (int r = 0; r < repeats; ++r) batch_launch_class(bt, ts, X, Y, sh, a, st);
                    HIP_OK(hipEventRecord(bt->ev1, st));
                    HIP_OK(hipEventSynchronize(bt->ev1));
                    float ms = 0.0f;
                    HIP_OK(hipEventElapsedTime(&ms, bt->ev0, bt->ev1));
                    const double per = ms / repeats;
                    if (block == 256 && wpc == 0) default_ms = per;
                    if (per < best_ms) {
                        best_ms = per;
                        best = sh;
                    }
                }
            bt->shape[X][Y] = best;
            bt->cost[X][Y] = best_ms;
            static const char* names[6] = {"ss", "ps", "pp", "ds", "dp", "dd"};
            char buf[256];
            std::snprintf(buf, sizeof(buf), "%s\"(%s|%s)\": {\"block\": %d, \"waves_per_cu\": %d, \"ms\": %.4f, "
                          "\"default_ms\": %.4f}", first ? "" : ", ", names[X], names[Y], best.block,
                          best.waves_per_cu, best_ms, default_ms);
            rep += buf;
            first = false;
        }
    rep += "}";
    bt->drop_graph();  // captured launches used the previous shapes
    return rep;
}

}  // namespace tlno_amd

extern "C" {

int tlno_gpu_batch_create(int n_mol, const int* n_comp, const double* centers, const int* ang, const int* nprim,
                          const double* exps, const double* coefs, tlno_gpu_batch** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("tlno_gpu: null output handle");
        *out = nullptr;
        if (n_mol <= 0 || !n_comp || !centers || !ang || !nprim || !exps || !coefs)
            throw std::runtime_error("eri_ao_tensor_hgp: empty or null batch.");
        auto bt = std::make_unique<tlno_gpu_batch>();
        bt->device = g_device;
        bt->nmol = n_mol;
        long long ci = 0, pi = 0;
        for (int m = 0; m < n_mol; ++m) {
            if (n_comp[m] <= 0) throw std::runtime_error("eri_ao_tensor_hgp: empty or null basis.");
            auto pl = std::make_unique<tlno_gpu_plan>();
            pl->device = g_device;
            pl->spherical = 0;
            long long np = 0;
            for (int i = 0; i < n_comp[m]; ++i) np += nprim[ci + i];
            build_plan(pl.get(), n_comp[m], centers + 3 * ci, ang + 3 * ci, nprim + ci, exps + pi, coefs + pi);
            if (pl->cls_begin[kNumLowClasses] != pl->npair)
                throw std::runtime_error("tlno_gpu: molecule batches take shells up to d in this build");
            ci += n_comp[m];
            pi += np;
            bt->mols.push_back(std::move(pl));
        }
        HIP_OK(hipSetDevice(bt->device));
        HIP_OK(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&bt->ev0));
        HIP_OK(hipEventCreate(&bt->ev1));
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, bt->device));
        bt->n_cu = std::max(1, prop.multiProcessorCount);
        batch_build_tables(bt.get());
        *out = bt.release();
    });
}

void tlno_gpu_batch_destroy(tlno_gpu_batch* batch) { delete batch; }

int tlno_gpu_batch_info(const tlno_gpu_batch* bt, int* n_mol, long long* total, int* M, long long* offsets) {
    return guarded([&] {
        if (!bt) throw std::runtime_error("tlno_gpu: null batch");
        if (n_mol) *n_mol = bt->nmol;
        if (total) *total = bt->total;
        for (int m = 0; m < bt->nmol; ++m) {
            if (M) M[m] = bt->mM[m];
            if (offsets) offsets[m] = bt->mbase[m];
        }
    });
}

int tlno_gpu_batch_build_jk_device(tlno_gpu_batch* bt, const double* d_density, double* d_J, double* d_K,
                                   double schwarz, int use_graph) {
    return guarded([&] {
        if (!bt) throw std::runtime_error("tlno_gpu: null batch");
        if (!d_density) throw std::runtime_error("eri_ao_tensor_hgp: null density.");
        batch_build(bt, d_density, d_J, d_K, schwarz, use_graph != 0);
    });
}

int tlno_gpu_batch_build_jk(tlno_gpu_batch* bt, const double* density, double* J, double* K, double schwarz) {
    return guarded([&] {
        if (!bt) throw std::runtime_error("tlno_gpu: null batch");
        if (!density) throw std::runtime_error("eri_ao_tensor_hgp: null density.");
        for (int m = 0; m < bt->nmol; ++m) validate_density(density + bt->mbase[m], bt->mM[m]);
        const std::size_t bytes = static_cast<std::size_t>(bt->total) * sizeof(double);
        double *dD = nullptr, *dJ = nullptr, *dK = nullptr;
        HIP_OK(hipSetDevice(bt->device));
        HIP_OK(hipMalloc(&dD, bytes));
        if (J) HIP_OK(hipMalloc(&dJ, bytes));
        if (K) HIP_OK(hipMalloc(&dK, bytes));
        HIP_OK(hipMemcpy(dD, density, bytes, hipMemcpyHostToDevice));
        batch_build(bt, dD, dJ, dK, schwarz, false);
        if (J) HIP_OK(hipMemcpy(J, dJ, bytes, hipMemcpyDeviceToHost));
        if (K) HIP_OK(hipMemcpy(K, dK, bytes, hipMemcpyDeviceToHost));
        HIP_OK(hipFree(dD));
        if (dJ) HIP_OK(hipFree(dJ));
        if (dK) HIP_OK(hipFree(dK));
    });
}

int tlno_gpu_batch_autotune(tlno_gpu_batch* bt, double schwarz, int repeats, char* report, int report_len) {
    return guarded([&] {
        if (!bt) throw std::runtime_error("tlno_gpu: null batch");
        if (!std::isfinite(schwarz) || schwarz < 0.0)
            throw std::runtime_error("eri_ao_tensor_hgp: schwarz_threshold must be finite and >= 0.");
        const std::string r = batch_autotune(bt, schwarz, std::max(1, repeats));
        if (report && report_len > 0) {
            std::strncpy(report, r.c_str(), static_cast<std::size_t>(report_len) - 1);
            report[report_len - 1] = '\0';
        }
    });
}

int tlno_gpu_batch_set_shape(tlno_gpu_batch* bt, int X, int Y, int block, int waves_per_cu) {
    return guarded([&] {
        if (!bt) throw std::runtime_error("tlno_gpu: null batch");
        if (X < 0 || X > 5 || Y < 0 || Y > X) throw std::runtime_error("tlno_gpu: class pair out of range");
        if (block != 64 && block != 128 && block != 256) throw std::runtime_error("tlno_gpu: block must be 64, 128 or 256");
        if (waves_per_cu < 0 || waves_per_cu > 64) throw std::runtime_error("tlno_gpu: waves per CU must be 0..64");
        bt->shape[X][Y] = LaunchShape{block, waves_per_cu};
        bt->drop_graph();
    });
}

int tlno_gpu_batch_last_stats(const tlno_gpu_batch* bt, long long* shell_quartets, long long* prim_quartets,
                              double* ms) {
    return guarded([&] {
        if (!bt) throw std::runtime_error("tlno_gpu: null batch");
        if (shell_quartets) *shell_quartets = bt->last_sq;
        if (prim_quartets) *prim_quartets = bt->last_pq;
        if (ms) *ms = bt->last_ms;
    });
}

}  // extern "C"
