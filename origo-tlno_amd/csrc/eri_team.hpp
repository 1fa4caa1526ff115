// Team evaluator for high angular momentum classes: one wave (four or six for
// the largest f and g classes, TeamShape::TW) per shell quartet, with the
// horizontal recurrence moved after the contraction.
//
// For bra shells (LA, LB) and ket shells (LC, LD) the contracted quartet is
//   (ab|cd) = HRR_bra HRR_ket [e|f],   [e|f] = sum_prim sum_root Gx Gy (w Gz),
// with e over the Cartesian triples LA <= |e| <= LA+LB, f over LC <= |f| <= LC+LD
// and G the 2D Rys tables of each axis (G[ex][fx], ex <= LA+LB, fx <= LC+LD).
// The HRR coefficients (AB, CD) do not depend on the primitives, so the
// transfer runs once per shell quartet instead of once per primitive root:
//   phase A  each lane takes one (ket primitive, bra primitive, root) combination
//            of a batch, evaluates its root and the three G tables and writes
//            them to LDS;
//   phase B  each lane owns one bra intermediate e and all [e|f] (the three
//            table rows of e are one read each per combination) and takes
//            every NP-th combination of the batch; the NP parts are summed
//            with cross-lane shuffles after the contraction;
//   phase C  [e|f] goes to LDS (to a per-workgroup global slot for the g
//            classes whose working set exceeds 144 KB), each lane expands its slice of the outputs by
//            the binomial form of the HRR, (a, b| = sum_{t <= b} C(b, t) AB^(b-t) [a+t|
//            (per axis), the same on the ket;
//   phase D  the block sits in LDS; lanes reduce J and K blocks with the
//            density blocks of the six shell pairs (one atomic per element).
// The digestion contract is that of the lane evaluator (eri_jk_task).
#pragma once

#include "eri_device.hpp"

namespace tlno_amd {

constexpr int ncart_cum(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }  // sum_{l <= L} ncart(l)

// index of triple (x, y, z), |e| = x + y + z, in the list of all triples with
// LMIN <= |e| ordered by |e| then descending lex
__host__ __device__ constexpr int tri_index(int LMIN, int x, int y, int z) {
    const int L = x + y + z;
    const int dx = L - x;
    return (ncart_cum(L - 1) - (LMIN > 0 ? ncart_cum(LMIN - 1) : 0)) + dx * (dx + 1) / 2 + (dx - y);
}

#ifndef TLNO_TEAM_LDS_KB
#define TLNO_TEAM_LDS_KB 20
#endif

template <int LA, int LB, int LC, int LD>
struct TeamShape {
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int NR = (LAB + LCD) / 2 + 1;
    static constexpr int NE = ncart_cum(LAB) - (LA > 0 ? ncart_cum(LA - 1) : 0);
    static constexpr int NF = ncart_cum(LCD) - (LC > 0 ? ncart_cum(LC - 1) : 0);
    static constexpr int NI = NE * NF;                     // [e|f] intermediates
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NOUT = NA * NB * NC * ND;
    // ket intermediates one thread holds through the contraction: up to 80
    // in registers; the g kets (gf| and (gg| (100, 145) split into two
    // f-chunks held by different waves
    static constexpr int NCH = (NF + 79) / 80;
    static constexpr int FCH = (NF + NCH - 1) / NCH;
    static_assert(NCH <= 2, "at most two f-chunks");
    static constexpr int EPAD = 64 * ((NE + 63) / 64);
    // waves per shell quartet: one thread per (bra intermediate, f-chunk) for
    // the f and g classes whose bra intermediates outnumber a wave's lanes
    // ((ff| has 74, (gg| 145), whose kets need f-chunks, or whose block
    // exceeds 3,000 values -- their HRR and digestion, over up to 50,625
    // values, dominate (f and g shells carry few primitives), and their LDS
    // leaves one workgroup per CU; at least four waves there (the first wave
    // alone evaluates the roots and tables, so classes of 1,800-2,200 values
    // with long contractions, (fd|fp) and (fd|dd), measured 35-50 % slower on
    // four waves and keep one); one wave for every class up to d
    static constexpr int TW = (NE > 64 || NOUT > 3000 || NCH > 1)
                                  ? (EPAD / 64 * NCH > 4 ? EPAD / 64 * NCH : 4) : 1;
    static constexpr int NT = 64 * TW;                     // threads per quartet
    // phase B ownership: thread = (bra intermediate e < EP, part); a thread
    // holds its f-chunk of the intermediates [e|f] of its e and takes every
    // NP-th combination (several waves: thread = (f-chunk, e), every
    // combination).  One wave: blocks of EP = NE lanes (EXACT) where that
    // keeps more lanes busy than blocks of the next power of two ((pp| has 9
    // intermediates: 7 parts on 63 lanes, where blocks of 16 left 7 of every
    // 16 idle); their partial rows are summed through LDS
    static constexpr int EP2 = NE <= 8 ? 8 : (NE <= 16 ? 16 : (NE <= 32 ? 32 : 64));
    static constexpr bool EXACT = TW == 1 && 64 / EP2 < 64 / NE;
    static constexpr int EP = TW > 1 ? EPAD : (EXACT ? NE : EP2);
    static constexpr int NP = TW > 1 ? 1 : 64 / EP;
    // every bra intermediate has its thread: one wave takes at most 64, the
    // multi-wave classes EPAD
    static_assert(TW > 1 || NE <= 64, "one wave owns at most 64 bra intermediates");
    static constexpr int OPL = (NOUT + 63) / 64;
    static constexpr int GS = (LAB + 1) * (LCD + 1);       // one axis table
    // one combination's three tables, padded to an odd number of doubles:
    // phase A has every lane write the same table entry of its own
    // combination, CS doubles apart, and an even stride maps lanes onto the
    // same LDS banks ((pp|pp), (dd|ds), (dd|dd) lost 8-16 % of their time to
    // bank conflicts with 3 GS + 1 even)
    static constexpr int CS = (3 * GS) | 1;
    // combinations per batch: 64 (one per lane) while their tables fit in
    // TLNO_TEAM_LDS_KB of LDS, else 32 (the upper half-wave then evaluates
    // the z tables of the lower half's combinations), so that two waves per
    // SIMD keep fitting in LDS
    static constexpr int CB = 64 * CS * 8 <= TLNO_TEAM_LDS_KB * 1024 ? 64 : 32;
    static constexpr int LDS_A = CB * CS;
    // phase C: [e|f] then the block share [0, R0); the ket-transformed
    // [e|cd] sits after them; phase D: the block, then six density blocks
    static constexpr int NX = NE * NC * ND;
    static constexpr int R0 = NI > NOUT ? NI : NOUT;
    // after [e|cd]: the unit-diagonal norms of the bra pairs (a, b) and of
    // the ket pairs (c, d), once per quartet
    static constexpr int LDS_C = R0 + NX + NA * NB + NC * ND;
    static constexpr int LDS_D = NOUT + NA * NB + NC * ND + NA * NC + NA * ND + NB * NC + NB * ND;
    static constexpr int LDS_CD = LDS_C > LDS_D ? LDS_C : LDS_D;
    // phases C and D work in LDS while their working set fits (gfx950 gives a
    // workgroup up to 160 KiB; (ff|ff) needs 138 KiB), else in a global
    // scratch slot of the workgroup (the large g classes: (gg|gg) needs 654 KiB)
    static constexpr bool GWORK = LDS_CD * 8 > 144 * 1024;
    static constexpr int WORK = GWORK ? LDS_CD : 0;        // doubles per scratch slot
    // the parts' partial rows [p][e][f] after [e|f] (EXACT blocks)
    static constexpr int LDS_P = EXACT ? (NP + 1) * NI : 0;
    static constexpr int LDS_ACD = LDS_A > LDS_CD ? LDS_A : LDS_CD;
    static constexpr int LDS = GWORK ? LDS_A : (LDS_ACD > LDS_P ? LDS_ACD : LDS_P);
    static_assert(LDS * 8 <= 160 * 1024, "team evaluator LDS beyond 160 KB");
    // waves per SIMD the register allocation must leave room for: one-wave
    // classes whose LDS admits eleven or more workgroups per CU ask for three
    // (at most 168 VGPRs); above that their registers, not their LDS, would
    // cap the CU at eight waves ((pp|pp), (dd|pp) and (dp|ps) at 176-186
    // VGPRs ran 18-27 % slower than at 164).  Classes that fit nine or ten
    // workgroups gain no wave from the cap and measured a few per cent slower
    // under it.
    static constexpr int WPE = (TW == 1 && 160 * 1024 / (LDS * 8) >= 11) ? 3 : 1;
    // phase-B unroll: four combinations per iteration where LDS, not
    // registers, caps the wave count and the rows are short ((dp|ds) 2.5 %
    // faster); two elsewhere (four cost (dd|ds) its third wave per SIMD)
    static constexpr int BU = (TW == 1 && WPE == 1 && NF <= 9) ? 4 : 2;
};


// binomial HRR weight C(n, t) X^(n - t) for n <= L (the shell's angular
// momentum, at most 4; the power n - t decides the coefficient: C(n, n-1) = n,
// C(n, n-2) = n(n-1)/2, C(n, n-3) = 4 for n = 4 and 1 for n = 3); exponents
// are decoded arithmetically, since a lane-dependent index into a constexpr
// table would put the table in per-lane scratch memory.  Shells up to d take
// the two-branch form (the general one costs the (pp|pp) quartets, whose
// transfer and digestion weigh as much as their contraction, about 25 %).
template <int L>
__device__ __forceinline__ double hrr_weight(int n, int t, double X) {
    static_assert(L <= 4, "HRR weights for shells up to g");
    const int k = n - t;
    if constexpr (L <= 2) {
        return k == 0 ? 1.0 : (k == 1 ? n * X : X * X);
    } else {
        if (k == 0) return 1.0;
        if (k == 1) return n * X;
        if (k == 2) return ((n * (n - 1)) >> 1) * (X * X);
        if (k == 3) return (n == 4 ? 4.0 : 1.0) * (X * X * X);
        return (X * X) * (X * X);
    }
}

// the triples of the list (|f| from LMIN to LMAX, by |f| then descending
// lex), as compile-time tables for unrolled loops
template <int LMIN, int LMAX>
struct TriTab {
    static constexpr int N = ncart_cum(LMAX) - (LMIN > 0 ? ncart_cum(LMIN - 1) : 0);
    int x[N], y[N], z[N];
    constexpr TriTab() : x(), y(), z() {
        int p = 0;
        for (int L = LMIN; L <= LMAX; ++L)
            for (int dx = 0; dx <= L; ++dx)
                for (int k = 0; k <= dx; ++k) {
                    x[p] = L - dx;
                    y[p] = dx - k;
                    z[p] = L - x[p] - y[p];
                    ++p;
                }
    }
};

// acc[f - F0] += Gx[ex][fx] Gy[ey][fy] Gz[ez][fz] for the ket intermediates
// f in [F0, F1) (compile-time triples, register accumulators)
template <int LC, int LD, int F0, int F1>
__device__ __forceinline__ void team_contract(const double* rx, const double* ry, const double* rz, double* acc) {
    constexpr TriTab<LC, LC + LD> ft{};
#pragma unroll
    for (int f = F0; f < F1; ++f) acc[f - F0] = fma(rx[ft.x[f]] * ry[ft.y[f]], rz[ft.z[f]], acc[f - F0]);
}

// triple of list position p (|e| from LMIN), decoded arithmetically
__device__ inline void tri_decode(int LMIN, int p, int& x, int& y, int& z) {
    int L = LMIN, base = 0;
    while (p >= base + ncart(L)) {
        base += ncart(L);
        ++L;
    }
    int k = p - base, dx = 0;
    while (k >= dx + 1) {
        k -= dx + 1;
        ++dx;
    }
    x = L - dx;
    y = dx - k;
    z = L - x - y;
}

// Phases A-C of one shell quartet, or of slice `slice` of ns of its
// combinations: leaves the unit-diagonal Cartesian block [a][b][c][d] in
// lds[0, NOUT) and returns true; returns false (nothing written) when the
// slice holds no combination.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ bool eri_team_block(const PairData& pd, const DevTables& tb, double fs_min,
                                               double* __restrict__ lds, double* __restrict__ work, int P, int Q,
                                               int slice, int ns) {
    using T = TeamShape<LA, LB, LC, LD>;
    const int tid = threadIdx.x;  // 0 .. NT - 1
    const int lane = tid & 63;

    // owned intermediates [e|f] for the f of the thread's chunk: their
    // factors are the rows Gx[ex][*], Gy[ey][*], Gz[ez][*] of a combination's
    // tables; the f-chunk is uniform over a wave (EP is a multiple of 64)
    const int part = T::TW > 1 ? 0 : tid / T::EP;
    const int chunk = T::TW > 1 ? tid / T::EP : 0;

    const double Ax = pd.A[3 * P], Ay = pd.A[3 * P + 1], Az = pd.A[3 * P + 2];
    const double Cx = pd.A[3 * Q], Cy = pd.A[3 * Q + 1], Cz = pd.A[3 * Q + 2];
    const int pb0 = pd.poff[P], pbn = pd.pn[P];
    const int pk0 = pd.poff[Q], pkn = pd.pn[Q];
    // slice `slice` of ns of the (ket primitive, bra primitive, root)
    // combinations, in whole batches; the block is linear in them, so the
    // slices of a quartet digest independently
    const int ntot = pkn * pbn * T::NR;
    const int per = ((ntot + ns - 1) / ns + T::CB - 1) / T::CB * T::CB;
    const int cbeg = slice * per;
    const int ncomb = min(ntot, cbeg + per);
    if (cbeg >= ncomb) return false;  // the whole block leaves together

    double* const tab = lds;  // phase A tables
    double* ef = work;        // NI, [e][f]
    const int e_own = tid % T::EP;
    int ex, ey, ez;
    tri_decode(LA, e_own < T::NE ? e_own : 0, ex, ey, ez);
    const int rx0 = ex * (T::LCD + 1), ry0 = T::GS + ey * (T::LCD + 1), rz0 = 2 * T::GS + ez * (T::LCD + 1);
    double acc[T::FCH];
#pragma unroll
    for (int f = 0; f < T::FCH; ++f) acc[f] = 0.0;
    for (int c0 = cbeg; c0 < ncomb; c0 += T::CB) {
        // ---- phase A: one combination per lane of the first wave; with
        // 32-combination batches the upper half of the wave evaluates the z
        // tables of the lower half's combinations (each half repeats the roots)
        const int cl = T::CB == 64 ? lane : (lane & 31);
        const int ax0 = T::CB == 64 ? 0 : (lane < 32 ? 0 : 2);
        const int ax1 = T::CB == 64 ? 3 : (lane < 32 ? 2 : 3);
        const int c = c0 + cl;
        if (c < ncomb && tid < 64) {
            const int r = c % T::NR;
            const int pq = c / T::NR;
            const int ib = pq % pbn, ik = pq / pbn;
            const double4 Pp = pd.pP[pb0 + ib];
            const double4 Qp = pd.pP[pk0 + ik];
            const double p = Pp.w, q = Qp.w;
            const double sipq = rsqrt_nr(p + q);
            const double ipq = sipq * sipq;
            const double PQx = Pp.x - Qp.x, PQy = Pp.y - Qp.y, PQz = Pp.z - Qp.z;
            const double Tv = p * q * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
            double pref = kTwoPi52 * pd.pw[pb0 + ib] * pd.pw[pk0 + ik] * sipq;
            if (fs_min > 0.0) {  // primitive FASTSCHWARZ screen: the combination adds zero
                const double w2 = pd.pw[pb0 + ib] * pd.pw[pk0 + ik], w4 = (w2 * w2) * (w2 * w2);
                if (w4 < (fs_min * fs_min) * (p * q)) pref = 0.0;
            }
            const RootEval<T::NR> re(tb, Tv);
            double t2, wr;
            re.get(tb, r, t2, wr);
            const double qo = q * ipq, po = p * ipq;
            const double B00 = 0.5 * ipq * t2;
            const double B10 = (0.5 * rcp_nr(p)) * (1.0 - qo * t2);
            const double B01 = (0.5 * rcp_nr(q)) * (1.0 - po * t2);
            const double qt = qo * t2, pt = po * t2;
            double* g = tab + cl * T::CS;
            // one axis at a time (x, y, z), so a single 2D table is live
#pragma unroll 1
            for (int ax = ax0; ax < ax1; ++ax) {
                const double c00 = ax == 0 ? Pp.x - Ax - qt * PQx : (ax == 1 ? Pp.y - Ay - qt * PQy : Pp.z - Az - qt * PQz);
                const double d00 = ax == 0 ? Qp.x - Cx + pt * PQx : (ax == 1 ? Qp.y - Cy + pt * PQy : Qp.z - Cz + pt * PQz);
                double G[T::LAB + 1][T::LCD + 1];
                G[0][0] = ax == 2 ? pref * wr : 1.0;
                if constexpr (T::LAB > 0) G[1][0] = c00 * G[0][0];
#pragma unroll
                for (int e = 1; e < T::LAB; ++e) G[e + 1][0] = c00 * G[e][0] + (e * B10) * G[e - 1][0];
#pragma unroll
                for (int f = 0; f < T::LCD; ++f)
#pragma unroll
                    for (int e = 0; e <= T::LAB; ++e) {
                        double v = d00 * G[e][f];
                        if (e > 0) v = fma(e * B00, G[e - 1][f], v);
                        if (f > 0) v = fma(f * B01, G[e][f - 1], v);
                        G[e][f + 1] = v;
                    }
                double* ga = g + ax * T::GS;
#pragma unroll
                for (int e = 0; e <= T::LAB; ++e)
#pragma unroll
                    for (int f = 0; f <= T::LCD; ++f) ga[e * (T::LCD + 1) + f] = G[e][f];
            }
        }
        __syncthreads();
        // ---- phase B: accumulate the owned intermediates over this part's
        // combinations of the batch (three LDS rows per combination)
        const int nb = (T::TW == 1 ? part < T::NP : chunk < T::NCH) ? min(T::CB, ncomb - c0) : 0;
#pragma unroll T::BU
        for (int j = part; j < nb; j += T::NP) {
            const double* g = tab + j * T::CS;
            double rx[T::LCD + 1], ry[T::LCD + 1], rz[T::LCD + 1];
#pragma unroll
            for (int t = 0; t <= T::LCD; ++t) {
                rx[t] = g[rx0 + t];
                ry[t] = g[ry0 + t];
                rz[t] = g[rz0 + t];
            }
            if constexpr (T::NCH == 1)
                team_contract<LC, LD, 0, T::NF>(rx, ry, rz, acc);
            else if (chunk == 0)
                team_contract<LC, LD, 0, T::FCH>(rx, ry, rz, acc);
            else
                team_contract<LC, LD, T::FCH, T::NF>(rx, ry, rz, acc);
        }
        __syncthreads();
    }

    if constexpr (T::EXACT) {
        // the parts' rows go to LDS after [e|f] (the tables are consumed);
        // part 0 adds them in part order
        double* prow = ef + T::NI;
        if (part < T::NP) {
#pragma unroll
            for (int f = 0; f < T::NF; ++f) prow[(part * T::NE + e_own) * T::NF + f] = acc[f];
        }
        __syncthreads();
        if (part == 0) {
#pragma unroll
            for (int f = 0; f < T::NF; ++f) {
                double sum = prow[e_own * T::NF + f];
#pragma unroll
                for (int pp = 1; pp < T::NP; ++pp) sum += prow[(pp * T::NE + e_own) * T::NF + f];
                ef[e_own * T::NF + f] = sum;
            }
        }
    } else {
        // sum the parts (lane bits above EP) into part 0, which stores its row
#pragma unroll
        for (int f = 0; f < T::FCH; ++f) {
            double v = acc[f];
#pragma unroll
            for (int m = T::EP; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
            acc[f] = v;
        }
        if (part == 0 && e_own < T::NE && chunk < T::NCH) {
            const int f0 = chunk * T::FCH;
#pragma unroll
            for (int f = 0; f < T::FCH; ++f)
                if (f0 + f < T::NF) ef[e_own * T::NF + f0 + f] = acc[f];
        }
    }
    __syncthreads();

    // ---- phase C: HRR from [e|f] to the unit-diagonal Cartesian block, ket
    // side first ([e|f] -> [e|cd] for every e), then bra side
    double* xk = work + T::R0;    // NX, [e][c][d]
    double* blk = work;           // NOUT, [a][b][c][d] (over [e|f], read before)
    const double AB[3] = {pd.AB[3 * P], pd.AB[3 * P + 1], pd.AB[3 * P + 2]};
    const double CD[3] = {pd.AB[3 * Q], pd.AB[3 * Q + 1], pd.AB[3 * Q + 2]};

    // norms of the pairs: the block element (ab|cd) takes nab[ab] ncd[cd]
    double* nab = xk + T::NX;
    double* ncd = nab + T::NA * T::NB;
    {
        auto nang = [](int x, int y, int z) { return inv_sqrt_small(dfact2m1(x) * dfact2m1(y) * dfact2m1(z)); };
        for (int t = tid; t < T::NA * T::NB + T::NC * T::ND; t += T::NT) {
            int x0, y0, z0, x1, y1, z1;
            if (t < T::NA * T::NB) {
                tri_decode(LA, t / T::NB, x0, y0, z0);
                tri_decode(LB, t % T::NB, x1, y1, z1);
            } else {
                const int u = t - T::NA * T::NB;
                tri_decode(LC, u / T::ND, x0, y0, z0);
                tri_decode(LD, u % T::ND, x1, y1, z1);
            }
            nab[t] = nang(x0, y0, z0) * nang(x1, y1, z1);  // ncd[u] = nab[NA NB + u]
        }
    }

    // [e|c d) = sum_u C(d, u) CD^(d - u) [e|c + u]
    for (int i = tid; i < T::NX; i += T::NT) {
        const int d = i % T::ND, c = (i / T::ND) % T::NC, e = i / (T::ND * T::NC);
        int cxc, cyc, czc, dx, dy, dz;
        tri_decode(LC, c, cxc, cyc, czc);
        tri_decode(LD, d, dx, dy, dz);
        const double* row = ef + e * T::NF;
        double u = 0.0;
        for (int ux = 0; ux <= dx; ++ux)
            for (int uy = 0; uy <= dy; ++uy)
                for (int uz = 0; uz <= dz; ++uz)
                    u = fma(hrr_weight<LD>(dx, ux, CD[0]) * hrr_weight<LD>(dy, uy, CD[1]) * hrr_weight<LD>(dz, uz, CD[2]),
                            row[tri_index(LC, cxc + ux, cyc + uy, czc + uz)], u);
        xk[i] = u;
    }
    __syncthreads();
    // (a b| = sum_t C(b, t) AB^(b - t) (a + t|, then the unit-diagonal norms
    for (int o = tid; o < T::NOUT; o += T::NT) {
        const int cd_ = o % (T::NC * T::ND), b = (o / (T::NC * T::ND)) % T::NB, a0 = o / (T::NC * T::ND * T::NB);
        int axc, ayc, azc, bx, by, bz;
        tri_decode(LA, a0, axc, ayc, azc);
        tri_decode(LB, b, bx, by, bz);
        double v = 0.0;
        for (int tx = 0; tx <= bx; ++tx)
            for (int ty = 0; ty <= by; ++ty)
                for (int tz = 0; tz <= bz; ++tz)
                    v = fma(hrr_weight<LB>(bx, tx, AB[0]) * hrr_weight<LB>(by, ty, AB[1]) * hrr_weight<LB>(bz, tz, AB[2]),
                            xk[tri_index(LA, axc + tx, ayc + ty, azc + tz) * (T::NC * T::ND) + cd_], v);
        blk[o] = v * (nab[a0 * T::NB + b] * ncd[cd_]);
    }
    __syncthreads();
    return true;
}

// Phase D: digestion of the block in lds[0, NOUT) (J and K element sums),
// the contract of the lane evaluator (eri_jk_task).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void eri_team_digest(const JKArgs& a, double* __restrict__ work, int P, int Q,
                                                double* __restrict__ Jm, double* __restrict__ Km,
                                                const double* __restrict__ Dm, long M) {
    using T = TeamShape<LA, LB, LC, LD>;
    const int tid = threadIdx.x;
    const PairData& pd = a.pd;
    double* blk = work;
    double s = 1.0;
    if (pd.diag[P]) s *= 0.5;
    if (pd.diag[Q]) s *= 0.5;
    if (P == Q) s *= 0.5;
    const int A0 = pd.ao[2 * P], B0 = pd.ao[2 * P + 1];
    const int C0 = pd.ao[2 * Q], D0 = pd.ao[2 * Q + 1];
    constexpr int NA = T::NA, NB = T::NB, NC = T::NC, ND = T::ND;
    // the six density blocks of the quartet's shells, staged in LDS after the block
    double* dab = blk + T::NOUT;
    double* dcd = dab + NA * NB;
    double* dac = dcd + NC * ND;
    double* dad = dac + NA * NC;
    double* dbc = dad + NA * ND;
    double* dbd = dbc + NB * NC;
    for (int t = tid; t < NA * NB; t += T::NT) dab[t] = Dm[(A0 + t / NB) * M + B0 + t % NB];
    for (int t = tid; t < NC * ND; t += T::NT) dcd[t] = Dm[(C0 + t / ND) * M + D0 + t % ND];
    for (int t = tid; t < NA * NC; t += T::NT) dac[t] = Dm[(A0 + t / NC) * M + C0 + t % NC];
    for (int t = tid; t < NA * ND; t += T::NT) dad[t] = Dm[(A0 + t / ND) * M + D0 + t % ND];
    for (int t = tid; t < NB * NC; t += T::NT) dbc[t] = Dm[(B0 + t / NC) * M + C0 + t % NC];
    for (int t = tid; t < NB * ND; t += T::NT) dbd[t] = Dm[(B0 + t / ND) * M + D0 + t % ND];
    __syncthreads();
    auto B = [&](int ai, int bi, int ci, int di) { return blk[((ai * NB + bi) * NC + ci) * ND + di]; };
    if (a.do_j) {
        // J[a][b] += 2 s sum_cd (ab|cd) D[c][d];  J[c][d] += 2 s sum_ab (ab|cd) D[a][b]
        for (int t = tid; t < NA * NB + NC * ND; t += T::NT) {
            double v = 0.0;
            if (t < NA * NB) {
                const int ai = t / NB, bi = t % NB;
                #pragma unroll 1
                for (int ci = 0; ci < NC; ++ci)
                    for (int di = 0; di < ND; ++di) v = fma(B(ai, bi, ci, di), dcd[ci * ND + di], v);
                unsafeAtomicAdd(Jm + (A0 + ai) * M + B0 + bi, 2.0 * s * v);
            } else {
                const int u = t - NA * NB, ci = u / ND, di = u % ND;
                #pragma unroll 1
                for (int ai = 0; ai < NA; ++ai)
                    for (int bi = 0; bi < NB; ++bi) v = fma(B(ai, bi, ci, di), dab[ai * NB + bi], v);
                unsafeAtomicAdd(Jm + (C0 + ci) * M + D0 + di, 2.0 * s * v);
            }
        }
    }
    if (a.do_k) {
        // K[a][c] += s sum_bd v D[b][d]; K[a][d] += s sum_bc v D[b][c];
        // K[b][c] += s sum_ad v D[a][d]; K[b][d] += s sum_ac v D[a][c]
        constexpr int n0 = NA * NC, n1 = NA * ND, n2 = NB * NC, n3 = NB * ND;
        for (int t = tid; t < n0 + n1 + n2 + n3; t += T::NT) {
            double v = 0.0;
            if (t < n0) {
                const int ai = t / NC, ci = t % NC;
                #pragma unroll 1
                for (int bi = 0; bi < NB; ++bi)
                    for (int di = 0; di < ND; ++di) v = fma(B(ai, bi, ci, di), dbd[bi * ND + di], v);
                unsafeAtomicAdd(Km + (A0 + ai) * M + C0 + ci, s * v);
            } else if (t < n0 + n1) {
                const int u = t - n0, ai = u / ND, di = u % ND;
                #pragma unroll 1
                for (int bi = 0; bi < NB; ++bi)
                    for (int ci = 0; ci < NC; ++ci) v = fma(B(ai, bi, ci, di), dbc[bi * NC + ci], v);
                unsafeAtomicAdd(Km + (A0 + ai) * M + D0 + di, s * v);
            } else if (t < n0 + n1 + n2) {
                const int u = t - n0 - n1, bi = u / NC, ci = u % NC;
                #pragma unroll 1
                for (int ai = 0; ai < NA; ++ai)
                    for (int di = 0; di < ND; ++di) v = fma(B(ai, bi, ci, di), dad[ai * ND + di], v);
                unsafeAtomicAdd(Km + (B0 + bi) * M + C0 + ci, s * v);
            } else {
                const int u = t - n0 - n1 - n2, bi = u / ND, di = u % ND;
                #pragma unroll 1
                for (int ai = 0; ai < NA; ++ai)
                    for (int ci = 0; ci < NC; ++ci) v = fma(B(ai, bi, ci, di), dac[ai * NC + ci], v);
                unsafeAtomicAdd(Km + (B0 + bi) * M + D0 + di, s * v);
            }
        }
    }
    __syncthreads();  // the LDS block is reused by the next quartet
}

// One wave (= one block of 64 threads) per shell quartet and combination
// slice: work item w is slice w % S (S = team_slices) of quartet
// tquart[w / S] (quartet mode), else of ket (w / S) % 64 of task w / (64 S).
// Same task formats, counters, molecule offsets and persistent stride (over
// work items) as the lane evaluator; FP64 only.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__((TeamShape<LA, LB, LC, LD>::NT))
__attribute__((amdgpu_waves_per_eu(TeamShape<LA, LB, LC, LD>::WPE))) void k_eri_team(JKArgs a) {
    using T = TeamShape<LA, LB, LC, LD>;
    __shared__ double lds[T::LDS];
    double* const work = T::GWORK ? a.gwork + static_cast<long long>(blockIdx.x) * T::WORK : lds;
    const long long w0 = blockIdx.x;
    const int ns = a.team_slices > 1 ? a.team_slices : 1;
    const long long nwork = (a.tquart ? a.ntquart : static_cast<long long>(a.ntasks) * 64) * ns;
    const long long end = a.stride > 0 ? nwork : w0 + 1;
    const long long step = a.stride > 0 ? a.stride : 1;
    for (long long w = w0; w < end; w += step) {
        const int slice = static_cast<int>(w % ns);
        const long long item = w / ns;
        int P, Q;
        if (a.tquart) {
            if (item >= a.ntquart) return;
            const int2 pq = a.tquart[item];
            P = __builtin_amdgcn_readfirstlane(pq.x);
            Q = __builtin_amdgcn_readfirstlane(pq.y);
        } else {
            const int task = static_cast<int>(item >> 6), k = static_cast<int>(item & 63);
            if (task >= a.ntasks) return;
            const int4 t = a.tasks[task];
            P = __builtin_amdgcn_readfirstlane(t.x);
            int nq;
            if (a.qlist) {
                const long long off = static_cast<long long>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(t.y))) |
                                      (static_cast<long long>(__builtin_amdgcn_readfirstlane(t.z) >> 8) << 32);
                nq = __builtin_amdgcn_readfirstlane(t.z) & 255;
                if (k >= nq) continue;
                Q = __builtin_amdgcn_readfirstlane(a.qlist[off + k]);
            } else {
                if ((__builtin_amdgcn_readfirstlane(t.w) & 0xffff) != 0) continue;  // slice 0 takes all primitives
                const int qb = __builtin_amdgcn_readfirstlane(t.y);
                int qe = __builtin_amdgcn_readfirstlane(t.z);
                if (a.cut) qe = min(qe, a.cut[6 * P]);
                nq = min(qe - qb, 64);
                if (k >= nq) continue;
                Q = qb + k;
            }
        }
        // (quartet lists with screens serve the f and g classes only, whose
        // bra pair holds the f or g shell; the code stays out of the
        // lower classes' kernels)
        if constexpr (LA >= 3) if (a.tquart && (a.sblk || a.sdecay)) {
            // the routing kernel's density and decay rules, per quartet
            bool keep = true;
            if (a.sblk) {
                const long long ng = a.sng;
                const long long ga = a.spgA[P], gb = a.spgB[P], gc = a.spgA[Q], gd = a.spgB[Q];
                const double coul = 4.0 * fmax(a.sblk[ga * ng + gb], a.sblk[gc * ng + gd]);
                const double exch = fmax(fmax(a.sblk[gc * ng + ga], a.sblk[gd * ng + ga]),
                                         fmax(a.sblk[gc * ng + gb], a.sblk[gd * ng + gb]));
                const double fac = a.smode == 1 ? coul : (a.smode == 2 ? exch : fmax(coul, exch));
                keep = a.sQ[P] * a.sQ[Q] * fac >= a.stau;
            }
            if (a.sdecay) keep = keep && a.sdecay[P] * a.sdecay[Q] >= a.sdecay_thr;
            if (!keep) continue;
        }
        long M = a.M;
        const double* __restrict__ Dm = a.D;
        double* __restrict__ Jm = a.J;
        double* __restrict__ Km = a.K;
        if (a.pmol) {
            const int mol = __builtin_amdgcn_readfirstlane(a.pmol[P]);
            const long long base = a.mbase[mol];
            M = a.mM[mol];
            Dm += base;
            Jm += base;
            Km += base;
        }
        if (a.counter && threadIdx.x == 0 && slice == 0) {
            unsigned long long* cnt = a.counter + 4 * (item % kCounterSlots);
            atomicAdd(cnt, 1ull);
            atomicAdd(cnt + 1, static_cast<unsigned long long>(a.pd.pn[Q]) * a.pd.pn[P]);
        }
        if (eri_team_block<LA, LB, LC, LD>(a.pd, a.tb, a.fs_min, lds, work, P, Q, slice, ns))
            eri_team_digest<LA, LB, LC, LD>(a, work, P, Q, Jm, Km, Dm, M);
    }
}

// Full unit-diagonal Cartesian blocks [a][b][c][d] of a list of shell
// quartets (Schwarz diagonal blocks and audit blocks of the classes that
// only the team evaluator serves): one wave per quartet, striding over the
// list when the grid is smaller.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__((TeamShape<LA, LB, LC, LD>::NT)) void k_eri_team_store(StoreArgs a) {
    using T = TeamShape<LA, LB, LC, LD>;
    __shared__ double lds[T::LDS];
    double* const work = T::GWORK ? a.gwork + static_cast<long long>(blockIdx.x) * T::WORK : lds;
    for (long long i = blockIdx.x; i < a.n; i += gridDim.x) {
        const int2 pq = a.quartets[i];
        const int P = __builtin_amdgcn_readfirstlane(pq.x);
        const int Q = __builtin_amdgcn_readfirstlane(pq.y);
        eri_team_block<LA, LB, LC, LD>(a.pd, a.tb, 0.0, lds, work, P, Q, 0, 1);
        double* out = a.out + (a.off ? a.off[i] : i * a.stride);
        for (int o = threadIdx.x; o < T::NOUT; o += T::NT) out[o] = work[o];
        __syncthreads();  // the LDS block is reused by the next quartet
    }
}

}  // namespace tlno_amd
