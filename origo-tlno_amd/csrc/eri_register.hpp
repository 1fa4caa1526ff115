// Compile-time registration of the per-class ERI kernels.  Each class pair
// (bra X >= ket Y) gets one J/K kernel and one store kernel whose output
// chunk is blockIdx.y:
//   full block in registers when it holds at most 81 values (one chunk),
//   otherwise one chunk per bra component (AC), and when that is still above
//   81 values, one per (bra component, ket component) pair (AC, CC).
#pragma once

#include <utility>

#include "eri_device.hpp"

namespace tlno_amd {

constexpr int kClassL[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

template <int LA, int LB, int LC, int LD>
struct ChunkPolicy {
    static constexpr int NA = ncart(LA), NC = ncart(LC);
    static constexpr int nc = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    static constexpr int mode = nc <= 81 ? 0 : (nc / NA <= 81 ? 1 : 2);
    static constexpr int nchunk = mode == 0 ? 1 : (mode == 1 ? NA : NA * NC);
    static constexpr int ac(int i) { return mode == 0 ? -1 : (mode == 1 ? i : i / NC); }
    static constexpr int cc(int i) { return mode == 2 ? i % NC : -1; }
};

template <typename R, int LA, int LB, int LC, int LD, int... I>
__global__ __launch_bounds__(256) void k_eri_jk(JKArgs a) {
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    const int ch = blockIdx.y;
    ((ch == I ? (eri_jk_body<R, LA, LB, LC, LD, CP::ac(I), CP::cc(I)>(a), 0) : 0), ...);
}

template <int LA, int LB, int LC, int LD, int... I>
__global__ __launch_bounds__(256) void k_eri_store(StoreArgs a) {
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    const int ch = blockIdx.y;
    ((ch == I ? (eri_store_body<LA, LB, LC, LD, CP::ac(I), CP::cc(I)>(a), 0) : 0), ...);
}

template <int LA, int LB, int LC, int LD, int... I>
void reg_kernels(ClassKernels& k, std::integer_sequence<int, I...>) {
    k.jk = k_eri_jk<double, LA, LB, LC, LD, I...>;
    k.jk32 = k_eri_jk<float, LA, LB, LC, LD, I...>;
    k.store = k_eri_store<LA, LB, LC, LD, I...>;
}

template <int X, int Y>
void reg_class(ClassKernels (&tab)[6][6]) {
    constexpr int LA = kClassL[X][0], LB = kClassL[X][1], LC = kClassL[Y][0], LD = kClassL[Y][1];
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    ClassKernels& k = tab[X][Y];
    k.nchunk = CP::nchunk;
    reg_kernels<LA, LB, LC, LD>(k, std::make_integer_sequence<int, CP::nchunk>{});
}

}  // namespace tlno_amd
