// Compile-time registration of the per-class ERI kernels.  Each class pair
// (bra X >= ket Y) gets one kernel instance per output chunk:
//   full block in registers when it holds at most 81 values,
//   otherwise one instance per bra component (AC), and when that is still
//   above 81 values, per (bra component, ket component) pair (AC, CC).
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
};

template <int LA, int LB, int LC, int LD, int AC, int CC>
void reg_one(ClassKernels& k, int idx) {
    k.jk[idx] = k_eri_jk<LA, LB, LC, LD, AC, CC>;
    k.store[idx] = k_eri_store<LA, LB, LC, LD, AC, CC>;
}

template <int LA, int LB, int LC, int LD, int... I>
void reg_chunks(ClassKernels& k, std::integer_sequence<int, I...>) {
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    if constexpr (CP::mode == 1) {
        (reg_one<LA, LB, LC, LD, I, -1>(k, I), ...);
    } else {
        (reg_one<LA, LB, LC, LD, I / CP::NC, I % CP::NC>(k, I), ...);
    }
}

template <int X, int Y>
void reg_class(ClassKernels (&tab)[6][6]) {
    constexpr int LA = kClassL[X][0], LB = kClassL[X][1], LC = kClassL[Y][0], LD = kClassL[Y][1];
    using CP = ChunkPolicy<LA, LB, LC, LD>;
    ClassKernels& k = tab[X][Y];
    k.nchunk = CP::nchunk;
    if constexpr (CP::mode == 0) {
        reg_one<LA, LB, LC, LD, -1, -1>(k, 0);
    } else {
        reg_chunks<LA, LB, LC, LD>(k, std::make_integer_sequence<int, CP::nchunk>{});
    }
}

}  // namespace tlno_amd
