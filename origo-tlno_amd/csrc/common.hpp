// Error plumbing shared by the C ABI translation units: HIP status checks
// throw, and every extern "C" entry point converts exceptions into a nonzero
// status plus the message returned by tlno_gpu_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace tlno_amd {

extern thread_local std::string g_err;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw ::tlno_amd::HipError(std::string("hip: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "tlno_gpu: unknown error";
        return 1;
    }
}

inline unsigned grid_for(long long n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

}  // namespace tlno_amd
