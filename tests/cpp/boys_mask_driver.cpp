// Masked-lane contract of the facade's Boys seeds, as a reference caller
// relies on it (benchmarks/one_electron_engine/src/boys_mask_contract.cpp:
// inactive inputs may be NaN, xs may be null when no lane is active, every
// inactive output is written as 0, nothing outside the W x (n+1) block is
// touched).  For every width W, order n in {0..8, 12, 32}, lane rotation and
// mask, the masked result must equal the unmasked one on the active lanes.
// Prints "PASS <checked>" and the unmasked seeds of the first rotation
// ("seeds W n v0 v1 ...") for an independent check against the oracle.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "core/math/general_math.hpp"

namespace cm = tlno::core::math;

static const double kArgs[8] = {0.0, 0.005, 39.999999999, 40.0, 40.000000001, 43.99999999, 56.0, 1000.0};
static const double kP = 0.75;

static bool is_finite_bits(double x) {
    std::uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    return (u >> 52 & 0x7ff) != 0x7ff;
}

template <int W>
static long run(FILE* seeds) {
    long checked = 0;
    const int orders[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 12, 32};
    for (int n : orders) {
        for (int rot = 0; rot < 8; ++rot) {
            double safe[W];
            for (int l = 0; l < W; ++l) safe[l] = kArgs[(l + rot) % 8];
            std::vector<double> full(static_cast<std::size_t>(n + 1) * W);
            cm::boys_r000_all_batch<W>(n, safe, kP, full.data());
            if (rot == 0) {
                std::fprintf(seeds, "seeds %d %d", W, n);
                for (double v : full) std::fprintf(seeds, " %.17g", v);
                std::fprintf(seeds, "\n");
            }
            for (unsigned mask = 0; mask < (1u << W); ++mask) {
                double xs[W];
                for (int l = 0; l < W; ++l)
                    xs[l] = (mask >> l & 1u) ? safe[l] : std::numeric_limits<double>::quiet_NaN();
                const double sentinel = -7.5;
                std::vector<double> out(static_cast<std::size_t>(n + 1) * W + 2, sentinel);
                cm::boys_r000_all_batch_masked<W>(n, mask ? xs : nullptr, kP, out.data() + 1, mask);
                if (out.front() != sentinel || out.back() != sentinel)
                    throw std::runtime_error("write outside the output block");
                for (int m = 0; m <= n; ++m)
                    for (int l = 0; l < W; ++l) {
                        const double got = out[1 + m * W + l];
                        if (!(mask >> l & 1u)) {
                            if (got != 0.0) throw std::runtime_error("inactive lane output is not zero");
                        } else {
                            const double want = full[m * W + l];
                            if (!is_finite_bits(got) || std::fabs(got - want) > 2e-13 * (1.0 + std::fabs(want)))
                                throw std::runtime_error("masked seed differs from the unmasked one");
                        }
                        ++checked;
                    }
            }
        }
    }
    return checked;
}

int main(int argc, char** argv) {
    FILE* seeds = argc > 1 ? std::fopen(argv[1], "w") : stdout;
    if (!seeds) return 2;
    try {
        const long n = run<1>(seeds) + run<2>(seeds) + run<4>(seeds) + run<8>(seeds);
        std::printf("PASS %ld\n", n);
    } catch (const std::exception& e) {
        std::printf("FAIL %s\n", e.what());
        return 1;
    }
    if (seeds != stdout) std::fclose(seeds);
    return 0;
}
