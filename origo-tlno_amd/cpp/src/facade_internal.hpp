// Shared pieces of the C++ facade's translation units: the GPU plan handle
// with its lock, the status check, and the component-basis upload.
#ifndef TLNO_AMD_FACADE_INTERNAL_HPP
#define TLNO_AMD_FACADE_INTERNAL_HPP

#include <cmath>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "core/basis/types.hpp"
#include "tlno_gpu.h"

namespace tlno_amd_facade {

using tlno::core::basis::Shell;

inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(tlno_gpu_last_error());
}

inline bool finite_nonneg(double v) { return std::isfinite(v) && v >= 0.0; }

// A GPU plan plus the lock that serializes builds on it (a plan owns one set
// of device work buffers).
struct GpuPlan {
    tlno_gpu_plan* h = nullptr;
    int M = 0, n_groups = 0, n_pairs = 0;
    long long n_prim_pairs = 0;
    std::mutex mu;
    ~GpuPlan() {
        if (h) tlno_gpu_plan_destroy(h);
    }
};

inline std::shared_ptr<GpuPlan> make_gpu_plan(const std::vector<Shell>& basis, int spherical) {
    const int n = static_cast<int>(basis.size());
    std::vector<double> centers(3 * static_cast<std::size_t>(n)), exps, coefs;
    std::vector<int> ang(3 * static_cast<std::size_t>(n)), nprim(n);
    for (int i = 0; i < n; ++i) {
        for (int x = 0; x < 3; ++x) {
            centers[3 * i + x] = basis[i].center[x];
            ang[3 * i + x] = basis[i].ang[x];
        }
        nprim[i] = static_cast<int>(basis[i].primitives.size());
        for (const auto& p : basis[i].primitives) {
            exps.push_back(p.exponent);
            coefs.push_back(p.coefficient);
        }
    }
    check(tlno_gpu_init(0));
    auto g = std::make_shared<GpuPlan>();
    check(tlno_gpu_plan_create(n, centers.data(), ang.data(), nprim.data(), exps.data(), coefs.data(), spherical,
                               &g->h));
    check(tlno_gpu_plan_info(g->h, &g->M, &g->n_groups, &g->n_pairs, &g->n_prim_pairs));
    return g;
}

}  // namespace tlno_amd_facade

#endif
