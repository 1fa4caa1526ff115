// Compile-time registration of the per-class ERI kernels.  Each class pair
// (bra X >= ket Y) is split into output chunks of at most TLNO_CHUNK_MAXOUT
// values:
//   mode 0: the full block (one chunk) of at most TLNO_CHUNK_SINGLE_MAXOUT;
//   mode 1: one chunk per run of AN consecutive bra components a (as many
//           as fit TLNO_CHUNK_RUN_MAXOUT; the last run may be shorter);
//   mode 2: one chunk per bra component a and group of CN ket components c
//           (the largest CN dividing ncart(LC) that fits the limit).
// A chunk holds its outputs in registers for the whole primitive loop, so
// fewer, larger chunks repeat less root and recurrence work at the price of
// registers.  J/K kernels hold TLNO_CHUNK_GROUP chunks each
// (chunk = c0 + blockIdx.y); the store kernel holds all chunks (blockIdx.y).
#pragma once

#include <utility>

#include "eri_device.hpp"
#include "eri_team.hpp"

#ifndef TLNO_CHUNK_MAXOUT
#define TLNO_CHUNK_MAXOUT 108
#endif
// blocks of 55-108 values run as bra-component runs of at most 54: (dp|ds)'s
// and (dd|ps)'s 108 values held 423-430 VGPRs in one chunk (one wave per
// SIMD); water40's FP64 build 1391 -> 1362 ms, its screened build 454 ->
// 438 ms
#ifndef TLNO_CHUNK_SINGLE_MAXOUT
#define TLNO_CHUNK_SINGLE_MAXOUT 54
#endif
// runs of bra components up to 54 outputs: (dp|pp)'s 162 values as three
// chunks of 54 hold 255 VGPRs (two waves per SIMD) where two chunks of 81
// took 390 with spills (one wave); water40's (dp|pp) 206 -> 185 ms in the
// FP64 build and 81 -> 65 ms in the screened one.  (dd|ds) runs as six
// chunks of 36; on organic30 its lane kernel is slower that way, and the
// team evaluator serves it there.
#ifndef TLNO_CHUNK_RUN_MAXOUT
#define TLNO_CHUNK_RUN_MAXOUT 54
#endif
#ifndef TLNO_CHUNK_GROUP
#define TLNO_CHUNK_GROUP 2
#endif

namespace tlno_amd {

constexpr int kClassL[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

template <int LA, int LB, int LC, int LD>
struct ChunkPolicy {
    static constexpr int NA = ncart(LA), NC = ncart(LC);
    static constexpr int nc = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    static constexpr int per_c = nc / NA / NC;  // outputs per (a, c)
    static constexpr int mode = nc <= TLNO_CHUNK_SINGLE_MAXOUT ? 0 : (nc / NA <= TLNO_CHUNK_MAXOUT ? 1 : 2);
    static constexpr int pick_cn() {
        for (int g = NC; g > 1; --g)
            if (NC % g == 0 && per_c * g <= TLNO_CHUNK_MAXOUT) return g;
        return 1;
    }
    static constexpr int cn = mode == 2 ? pick_cn() : 1;
    static constexpr int ncg = NC / cn;  // ket component groups (mode 2)
    static constexpr int per_a = nc / NA;  // outputs per bra component
    static constexpr int an_max = TLNO_CHUNK_RUN_MAXOUT / per_a > 1 ? TLNO_CHUNK_RUN_MAXOUT / per_a : 1;
    static constexpr int an_run = mode == 1 ? (an_max < NA ? an_max : NA) : 1;
    static constexpr int nchunk = mode == 0 ? 1 : (mode == 1 ? (NA + an_run - 1) / an_run : NA * ncg);
    static constexpr int ac(int i) { return mode == 0 ? -1 : (mode == 1 ? i * an_run : i / ncg); }
    static constexpr int an(int i) {
        return mode == 1 ? (NA - i * an_run < an_run ? NA - i * an_run : an_run) : 1;
    }
    static constexpr int cc(int i) { return mode == 2 ? (i % ncg) * cn : -1; }
};

// Classes with one or two Rys roots take their roots from the Boys moments
// F_0..F_3, gathered per primitive quartet: the block stages the coarse Boys
// table (31.6 KB) in LDS once, and the gathers go to LDS instead of through
// the texture path, which bounds these classes.
template <typename R, int LA, int LB, int LC, int LD, bool FS, int C0, int... J>
__global__ __launch_bounds__(256) void k_eri_jk(JKArgs a) {
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    constexpr bool kLdsBoys = (LA + LB + LC + LD) / 2 + 1 <= 2;
    __shared__ double sboys[kLdsBoys ? BOYS_LDS_NGRID * BOYS_LDS_M : 1];
    const double* sb = nullptr;
    if constexpr (kLdsBoys) {
        for (int i = threadIdx.x; i < BOYS_LDS_NGRID * BOYS_LDS_M; i += blockDim.x) sboys[i] = a.tb.boys_lds[i];
        __syncthreads();
        sb = sboys;
    }
    const int ch = C0 + static_cast<int>(blockIdx.y);
    ((ch == C0 + J ? (eri_jk_body<R, LA, LB, LC, LD, CP::ac(C0 + J), CP::an(C0 + J), CP::cc(C0 + J), CP::cn, FS>(a, sb), 0)
                   : 0),
     ...);
}

template <int LA, int LB, int LC, int LD, int... I>
__global__ __launch_bounds__(256) void k_eri_store(StoreArgs a) {
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    const int ch = blockIdx.y;
    ((ch == I ? (eri_store_body<LA, LB, LC, LD, CP::ac(I), CP::an(I), CP::cc(I), CP::cn>(a), 0) : 0), ...);
}

template <int LA, int LB, int LC, int LD, int... I>
void reg_store(ClassKernels& k, std::integer_sequence<int, I...>) {
    k.store = k_eri_store<LA, LB, LC, LD, I...>;
}

template <int LA, int LB, int LC, int LD, int C0, int... J>
void reg_group(ClassKernels& k, std::integer_sequence<int, J...>) {
    JKGroup& g = k.grp[k.ngroup++];
    g.jk = k_eri_jk<double, LA, LB, LC, LD, false, C0, J...>;
    g.jk32 = k_eri_jk<float, LA, LB, LC, LD, false, C0, J...>;
    g.jkfs = k_eri_jk<double, LA, LB, LC, LD, true, C0, J...>;
    g.jk32fs = k_eri_jk<float, LA, LB, LC, LD, true, C0, J...>;
    g.c0 = C0;
    g.n = static_cast<int>(sizeof...(J));
}

template <int LA, int LB, int LC, int LD, int... G>
void reg_groups(ClassKernels& k, std::integer_sequence<int, G...>) {
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    constexpr int GS = TLNO_CHUNK_GROUP;
    (reg_group<LA, LB, LC, LD, G * GS>(
         k, std::make_integer_sequence<int, (CP::nchunk - G * GS < GS ? CP::nchunk - G * GS : GS)>{}),
     ...);
}

template <int X, int Y>
void reg_class(ClassKernels (&tab)[6][6]) {
    constexpr int LA = kClassL[X][0], LB = kClassL[X][1], LC = kClassL[Y][0], LD = kClassL[Y][1];
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    constexpr int ngroup = (CP::nchunk + TLNO_CHUNK_GROUP - 1) / TLNO_CHUNK_GROUP;
    static_assert(ngroup <= kMaxChunkGroups, "too many chunk groups");
    ClassKernels& k = tab[X][Y];
    k.nchunk = CP::nchunk;
    k.ngroup = 0;
    reg_groups<LA, LB, LC, LD>(k, std::make_integer_sequence<int, ngroup>{});
    reg_store<LA, LB, LC, LD>(k, std::make_integer_sequence<int, CP::nchunk>{});
    k.nout = CP::nc;
    if constexpr (LA + LB + LC + LD >= 4) {
        static_assert(TeamShape<LA, LB, LC, LD>::NT == 64, "classes up to d run one wave per quartet");
        k.team = k_eri_team<LA, LB, LC, LD>;
    }
}

// (la, lb) of every pair class, the f classes after the six of the lane evaluator
constexpr int kClassLAll[kNumClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}, {3, 0}, {3, 1},
                                            {3, 2}, {3, 3}, {4, 0}, {4, 1}, {4, 2}, {4, 3}, {4, 4}};

template <int X, int Y>
void reg_high(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    constexpr int LA = kClassLAll[X][0], LB = kClassLAll[X][1], LC = kClassLAll[Y][0], LD = kClassLAll[Y][1];
    HighKernels& k = tab[X - kNumLowClasses][Y];
    k.team = k_eri_team<LA, LB, LC, LD>;
    k.store = k_eri_team_store<LA, LB, LC, LD>;
    k.nout = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    k.threads = TeamShape<LA, LB, LC, LD>::NT;
    k.gwork = TeamShape<LA, LB, LC, LD>::WORK;
}

}  // namespace tlno_amd
