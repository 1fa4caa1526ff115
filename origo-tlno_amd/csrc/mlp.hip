// Training and inference of the learned-screen network (mlp.hpp) on the GPU:
// a fused forward + backward + weight-gradient kernel (one wave = 64 samples,
// per-sample vectors staged in LDS, block partial gradients added atomically),
// an Adam kernel that also refreshes the bf16 working copy of the weights,
// and two inference kernels (per-lane FP32 and wave-level bf16 MFMA).
//
// Loss: mean over samples of c(r) r^2 with r = t - y and c = w_under for
// r > 0 (under-predicted magnitude, the unsafe side of a screen), 1 otherwise.
// Gradients treat the bf16 roundings as identity (straight-through).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "mlp.hpp"
#include "tlno_gpu.h"

namespace tlno_amd {

__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ w, const float* __restrict__ X,
                                                 const float* __restrict__ T, const int* __restrict__ idx,
                                                 long long n, float inv_n, float w_under,
                                                 float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float sx[64][MLP_IN + 1];   // bf16(normalized x)
    __shared__ float sh1[64][MLP_H + 1];   // h1 (bf16 values)
    __shared__ float sd1[64][MLP_H + 1];   // z1, then dL/dz1
    __shared__ float sd2[64][MLP_H + 1];   // z2, then dL/dz2
    __shared__ float sh2[64][MLP_H + 1];   // h2
    __shared__ float sdy[64];
    const int s = threadIdx.x;
    const long long gi = static_cast<long long>(blockIdx.x) * 64 + s;
    const bool valid = gi < n;
    const long long row = valid ? (idx ? idx[gi] : gi) : 0;
    for (int i = 0; i < MLP_IN; ++i)
        sx[s][i] = bf16_round((X[row * MLP_IN + i] - w[MLP_OFF_MEAN + i]) * w[MLP_OFF_ISTD + i]);
    for (int j = 0; j < MLP_H; ++j) {
        float acc = 0.0f;
        for (int i = 0; i < MLP_IN; ++i) acc = fmaf(w[MLP_OFF_W1 + j * MLP_IN + i], sx[s][i], acc);
        const float z = acc + w[MLP_OFF_B1 + j];
        sd1[s][j] = z;
        sh1[s][j] = bf16_round(fmaxf(z, 0.0f));
    }
    float y = 0.0f;
    for (int j = 0; j < MLP_H; ++j) {
        float acc = 0.0f;
        for (int i = 0; i < MLP_H; ++i) acc = fmaf(w[MLP_OFF_W2 + j * MLP_H + i], sh1[s][i], acc);
        const float z = acc + w[MLP_OFF_B2 + j];
        sd2[s][j] = z;
        const float hv = fmaxf(z, 0.0f);
        sh2[s][j] = hv;
        y = fmaf(w[MLP_OFF_W3 + j], hv, y);
    }
    y += w[MLP_OFF_B3];
    const float r = valid ? T[row] - y : 0.0f;
    const float c = r > 0.0f ? w_under : 1.0f;
    const float dy = valid ? -2.0f * c * r * inv_n : 0.0f;
    float l = valid ? c * r * r * inv_n : 0.0f;
    sdy[s] = dy;
    for (int j = 0; j < MLP_H; ++j) sd2[s][j] = sd2[s][j] > 0.0f ? dy * w[MLP_OFF_W3 + j] : 0.0f;
    for (int i = 0; i < MLP_H; ++i) {
        float acc = 0.0f;
        for (int j = 0; j < MLP_H; ++j) acc = fmaf(w[MLP_OFF_W2 + j * MLP_H + i], sd2[s][j], acc);
        sd1[s][i] = sd1[s][i] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    for (int p = s; p < MLP_NPARAM; p += 64) {
        float acc = 0.0f;
        if (p < MLP_OFF_B1) {
            const int j = p / MLP_IN, i = p % MLP_IN;
            for (int k = 0; k < 64; ++k) acc = fmaf(sd1[k][j], sx[k][i], acc);
        } else if (p < MLP_OFF_W2) {
            const int j = p - MLP_OFF_B1;
            for (int k = 0; k < 64; ++k) acc += sd1[k][j];
        } else if (p < MLP_OFF_B2) {
            const int j = (p - MLP_OFF_W2) / MLP_H, i = (p - MLP_OFF_W2) % MLP_H;
            for (int k = 0; k < 64; ++k) acc = fmaf(sd2[k][j], sh1[k][i], acc);
        } else if (p < MLP_OFF_W3) {
            const int j = p - MLP_OFF_B2;
            for (int k = 0; k < 64; ++k) acc += sd2[k][j];
        } else if (p < MLP_OFF_B3) {
            const int j = p - MLP_OFF_W3;
            for (int k = 0; k < 64; ++k) acc = fmaf(sdy[k], sh2[k][j], acc);
        } else {
            for (int k = 0; k < 64; ++k) acc += sdy[k];
        }
        atomicAdd(grad + p, acc);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) l += __shfl_xor(l, m, 64);
    if (s == 0) atomicAdd(loss, l);
}

// Adam with bias correction (c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)); the
// gradient buffer is cleared for the next step and wb receives the bf16
// working copy (biases and normalization stay FP32).
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ wb, float* __restrict__ m,
                           float* __restrict__ v, float* __restrict__ g, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    wb[i] = mlp_is_weight(i) ? bf16_round(wi) : wi;
    g[i] = 0.0f;
}

__global__ void k_fill_f32(float* p, long long n, float v) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_mlp_predict_lane(const float* __restrict__ wb, const float* __restrict__ X, long long n,
                                   float* __restrict__ y) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[MLP_IN];
    for (int k = 0; k < MLP_IN; ++k) x[k] = X[i * MLP_IN + k];
    y[i] = mlp_forward_lane(wb, x);
}

__global__ __launch_bounds__(256) void k_mlp_predict_wave(MlpPacked mp, const float* __restrict__ X, long long n,
                                                          float* __restrict__ y) {
    const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const long long i = wave * 64 + (threadIdx.x & 63);
    if (wave * 64 >= n) return;  // whole wave leaves together
    const long long src = i < n ? i : n - 1;
    float x[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) x[k] = X[src * MLP_IN + k];
    const float v = mlp_forward_wave(mp, x);
    if (i < n) y[i] = v;
}

// host-side deterministic generator (splitmix64)
struct HostRng {
    uint64_t s;
    explicit HostRng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace tlno_amd

using namespace tlno_amd;

struct tlno_gpu_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<float> w;  // host mirror of the master parameters + normalization (MLP_NALL)
    float *d_w = nullptr, *d_wb = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr, *d_loss = nullptr;
    long long step = 0;
    // packed MFMA operands (refreshed by sync_packed)
    uint4* d_frag = nullptr;
    float* d_vec = nullptr;
    float* d_norm = nullptr;
    float b3 = 0.0f;
    ~tlno_gpu_mlp() {
        for (void* p : {(void*)d_w, (void*)d_wb, (void*)d_m, (void*)d_v, (void*)d_g, (void*)d_loss, (void*)d_frag,
                        (void*)d_vec, (void*)d_norm})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void upload_params(tlno_gpu_mlp* m) {
    std::vector<float> wb(m->w);
    for (int i = 0; i < MLP_NPARAM; ++i)
        if (mlp_is_weight(i)) wb[i] = bf16_round(wb[i]);
    HIP_OK(hipMemcpy(m->d_w, m->w.data(), MLP_NALL * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(m->d_wb, wb.data(), MLP_NALL * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(m->d_m, 0, MLP_NPARAM * sizeof(float)));
    HIP_OK(hipMemset(m->d_v, 0, MLP_NPARAM * sizeof(float)));
    HIP_OK(hipMemset(m->d_g, 0, MLP_NPARAM * sizeof(float)));
    m->step = 0;
}

void download_params(tlno_gpu_mlp* m) {
    HIP_OK(hipMemcpy(m->w.data(), m->d_w, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
}

void sync_packed(tlno_gpu_mlp* m) {
    std::vector<uint4> frag(MLP_NFRAG * 64);
    std::vector<float> vec(192), norm(64);
    mlp_pack(m->w.data(), frag.data(), vec.data(), norm.data(), &m->b3);
    HIP_OK(hipMemcpy(m->d_frag, frag.data(), frag.size() * sizeof(uint4), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(m->d_vec, vec.data(), vec.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(m->d_norm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice));
}

void check_mlp(const tlno_gpu_mlp* m) {
    if (!m) throw std::runtime_error("tlno_gpu: null network handle");
}

void check_data(long long n, const float* X, const float* t) {
    if (n <= 0) throw std::runtime_error("tlno_gpu: training set must not be empty");
    if (!X || !t) throw std::runtime_error("tlno_gpu: null training data");
    for (long long i = 0; i < n; ++i) {
        if (!std::isfinite(t[i])) throw std::runtime_error("tlno_gpu: non-finite training label");
        for (int k = 0; k < MLP_IN; ++k)
            if (!std::isfinite(X[i * MLP_IN + k])) throw std::runtime_error("tlno_gpu: non-finite feature");
    }
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(std::size_t bytes) { HIP_OK(hipMalloc(&p, std::max<std::size_t>(bytes, 16))); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// one gradient accumulation over samples idx[0..n) (or 0..n) of X/T into d_g and d_loss
void accumulate_grad(tlno_gpu_mlp* m, const float* dX, const float* dT, const int* didx, long long n,
                     float inv_n, float w_under) {
    const unsigned blocks = grid_for(n, 64);
    k_mlp_grad<<<blocks, 64, 0, m->stream>>>(m->d_wb, dX, dT, didx, n, inv_n, w_under, m->d_g, m->d_loss);
    HIP_OK(hipGetLastError());
}

void adam(tlno_gpu_mlp* m, float lr) {
    constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++m->step;
    const float c1 = static_cast<float>(1.0 / (1.0 - std::pow(double(b1), double(m->step))));
    const float c2 = static_cast<float>(1.0 / (1.0 - std::pow(double(b2), double(m->step))));
    k_mlp_adam<<<grid_for(MLP_NPARAM, 256), 256, 0, m->stream>>>(m->d_w, m->d_wb, m->d_m, m->d_v, m->d_g, lr, b1,
                                                                  b2, eps, c1, c2);
    HIP_OK(hipGetLastError());
}

}  // namespace

extern "C" {

int tlno_gpu_mlp_create(int device, unsigned seed, tlno_gpu_mlp** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("tlno_gpu: null output handle");
        *out = nullptr;
        std::unique_ptr<tlno_gpu_mlp> m(new tlno_gpu_mlp);
        m->device = device;
        HIP_OK(hipSetDevice(device));
        HIP_OK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        HIP_OK(hipMalloc(&m->d_w, MLP_NALL * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_wb, MLP_NALL * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_m, MLP_NPARAM * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_v, MLP_NPARAM * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_g, MLP_NPARAM * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_loss, 4 * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_frag, MLP_NFRAG * 64 * sizeof(uint4)));
        HIP_OK(hipMalloc(&m->d_vec, 192 * sizeof(float)));
        HIP_OK(hipMalloc(&m->d_norm, 64 * sizeof(float)));
        // He-uniform hidden layers, Glorot-uniform output row, zero biases,
        // identity normalization
        m->w.assign(MLP_NALL, 0.0f);
        HostRng rng(seed);
        const double a1 = std::sqrt(6.0 / MLP_IN), a2 = std::sqrt(6.0 / MLP_H), a3 = std::sqrt(6.0 / (MLP_H + 1));
        for (int i = 0; i < MLP_H * MLP_IN; ++i) m->w[MLP_OFF_W1 + i] = static_cast<float>(a1 * (2 * rng.uniform() - 1));
        for (int i = 0; i < MLP_H * MLP_H; ++i) m->w[MLP_OFF_W2 + i] = static_cast<float>(a2 * (2 * rng.uniform() - 1));
        for (int i = 0; i < MLP_H; ++i) m->w[MLP_OFF_W3 + i] = static_cast<float>(a3 * (2 * rng.uniform() - 1));
        for (int i = 0; i < MLP_IN; ++i) m->w[MLP_OFF_ISTD + i] = 1.0f;
        upload_params(m.get());
        *out = m.release();
    });
}

void tlno_gpu_mlp_destroy(tlno_gpu_mlp* m) { delete m; }

int tlno_gpu_mlp_get_params(tlno_gpu_mlp* m, float* out) {
    return guarded([&] {
        check_mlp(m);
        if (!out) throw std::runtime_error("tlno_gpu: null output");
        HIP_OK(hipSetDevice(m->device));
        HIP_OK(hipStreamSynchronize(m->stream));
        download_params(m);
        std::memcpy(out, m->w.data(), MLP_NALL * sizeof(float));
    });
}

int tlno_gpu_mlp_set_params(tlno_gpu_mlp* m, const float* in) {
    return guarded([&] {
        check_mlp(m);
        if (!in) throw std::runtime_error("tlno_gpu: null parameters");
        for (int i = 0; i < MLP_NALL; ++i)
            if (!std::isfinite(in[i])) throw std::runtime_error("tlno_gpu: non-finite network parameter");
        HIP_OK(hipSetDevice(m->device));
        HIP_OK(hipStreamSynchronize(m->stream));
        m->w.assign(in, in + MLP_NALL);
        upload_params(m);
    });
}

int tlno_gpu_mlp_set_normalization(tlno_gpu_mlp* m, long long n, const float* X) {
    return guarded([&] {
        check_mlp(m);
        if (n <= 0 || !X) throw std::runtime_error("tlno_gpu: normalization needs samples");
        HIP_OK(hipSetDevice(m->device));
        HIP_OK(hipStreamSynchronize(m->stream));
        download_params(m);
        for (int k = 0; k < MLP_IN; ++k) {
            double s = 0.0, s2 = 0.0;
            for (long long i = 0; i < n; ++i) {
                const double v = X[i * MLP_IN + k];
                s += v;
                s2 += v * v;
            }
            const double mean = s / n;
            const double var = std::max(0.0, s2 / n - mean * mean);
            m->w[MLP_OFF_MEAN + k] = static_cast<float>(mean);
            m->w[MLP_OFF_ISTD + k] = static_cast<float>(1.0 / std::max(std::sqrt(var), 1e-3));
        }
        upload_params(m);
    });
}

int tlno_gpu_mlp_loss_grad(tlno_gpu_mlp* m, long long n, const float* X, const float* t, float w_under,
                           float* grad, double* loss) {
    return guarded([&] {
        check_mlp(m);
        check_data(n, X, t);
        HIP_OK(hipSetDevice(m->device));
        DevBuf dX(n * MLP_IN * sizeof(float)), dT(n * sizeof(float));
        HIP_OK(hipMemcpy(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dT.p, t, n * sizeof(float), hipMemcpyHostToDevice));
        k_fill_f32<<<grid_for(MLP_NPARAM, 256), 256, 0, m->stream>>>(m->d_g, MLP_NPARAM, 0.0f);
        k_fill_f32<<<1, 64, 0, m->stream>>>(m->d_loss, 1, 0.0f);
        accumulate_grad(m, dX.as<float>(), dT.as<float>(), nullptr, n, static_cast<float>(1.0 / n), w_under);
        float l = 0.0f;
        std::vector<float> g(MLP_NPARAM);
        HIP_OK(hipMemcpyAsync(g.data(), m->d_g, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIP_OK(hipMemcpyAsync(&l, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIP_OK(hipStreamSynchronize(m->stream));
        k_fill_f32<<<grid_for(MLP_NPARAM, 256), 256, 0, m->stream>>>(m->d_g, MLP_NPARAM, 0.0f);
        HIP_OK(hipStreamSynchronize(m->stream));
        if (grad) std::memcpy(grad, g.data(), MLP_NPARAM * sizeof(float));
        if (loss) *loss = l;
    });
}

int tlno_gpu_mlp_adam_step(tlno_gpu_mlp* m, long long n, const float* X, const float* t, float lr, float w_under) {
    return guarded([&] {
        check_mlp(m);
        check_data(n, X, t);
        if (!(lr > 0.0f) || !std::isfinite(lr)) throw std::runtime_error("tlno_gpu: learning rate must be > 0");
        HIP_OK(hipSetDevice(m->device));
        DevBuf dX(n * MLP_IN * sizeof(float)), dT(n * sizeof(float));
        HIP_OK(hipMemcpy(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dT.p, t, n * sizeof(float), hipMemcpyHostToDevice));
        k_fill_f32<<<1, 64, 0, m->stream>>>(m->d_loss, 1, 0.0f);
        accumulate_grad(m, dX.as<float>(), dT.as<float>(), nullptr, n, static_cast<float>(1.0 / n), w_under);
        adam(m, lr);
        HIP_OK(hipStreamSynchronize(m->stream));
        download_params(m);
    });
}

int tlno_gpu_mlp_train(tlno_gpu_mlp* m, long long n, const float* X, const float* t, int epochs, int batch,
                       float lr, float w_under, unsigned seed, float* epoch_loss) {
    return guarded([&] {
        check_mlp(m);
        check_data(n, X, t);
        if (epochs <= 0 || batch <= 0) throw std::runtime_error("tlno_gpu: epochs and batch must be > 0");
        if (!(lr > 0.0f) || !std::isfinite(lr)) throw std::runtime_error("tlno_gpu: learning rate must be > 0");
        if (n > (1ll << 31) - 1) throw std::runtime_error("tlno_gpu: training set too large");
        HIP_OK(hipSetDevice(m->device));
        DevBuf dX(n * MLP_IN * sizeof(float)), dT(n * sizeof(float)), dI(n * sizeof(int));
        HIP_OK(hipMemcpy(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dT.p, t, n * sizeof(float), hipMemcpyHostToDevice));
        std::vector<int> perm(n);
        for (long long i = 0; i < n; ++i) perm[i] = static_cast<int>(i);
        HostRng rng(seed);
        const long long steps_per_epoch = (n + batch - 1) / batch;
        const long long total = steps_per_epoch * epochs;
        long long k = 0;
        for (int e = 0; e < epochs; ++e) {
            for (long long i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng.next() % (i + 1)]);
            HIP_OK(hipMemcpyAsync(dI.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, m->stream));
            k_fill_f32<<<1, 64, 0, m->stream>>>(m->d_loss, 1, 0.0f);
            for (long long b0 = 0; b0 < n; b0 += batch, ++k) {
                const long long nb = std::min<long long>(batch, n - b0);
                // cosine decay from lr to lr / 20
                const double f = 0.5 * (1.0 + std::cos(M_PI * double(k) / double(total)));
                const float lr_k = static_cast<float>(lr * (0.05 + 0.95 * f));
                // gradient of the batch-mean loss; the batch means add up in d_loss
                accumulate_grad(m, dX.as<float>(), dT.as<float>(), dI.as<int>() + b0, nb,
                                static_cast<float>(1.0 / nb), w_under);
                adam(m, lr_k);
            }
            float l = 0.0f;
            HIP_OK(hipMemcpyAsync(&l, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
            HIP_OK(hipStreamSynchronize(m->stream));
            if (epoch_loss) epoch_loss[e] = l / static_cast<float>(steps_per_epoch);
        }
        HIP_OK(hipStreamSynchronize(m->stream));
        download_params(m);
    });
}

int tlno_gpu_mlp_predict(tlno_gpu_mlp* m, long long n, const float* X, float* y, int matrix_cores) {
    return guarded([&] {
        check_mlp(m);
        if (n < 0 || (n > 0 && (!X || !y))) throw std::runtime_error("tlno_gpu: bad prediction arguments");
        if (n == 0) return;
        HIP_OK(hipSetDevice(m->device));
        HIP_OK(hipStreamSynchronize(m->stream));
        DevBuf dX(n * MLP_IN * sizeof(float)), dY(n * sizeof(float));
        HIP_OK(hipMemcpy(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        if (matrix_cores) {
            download_params(m);
            sync_packed(m);
            MlpPacked mp{m->d_frag, m->d_vec, m->d_norm, m->b3};
            k_mlp_predict_wave<<<grid_for(n, 256), 256, 0, m->stream>>>(mp, dX.as<float>(), n, dY.as<float>());
        } else {
            k_mlp_predict_lane<<<grid_for(n, 256), 256, 0, m->stream>>>(m->d_wb, dX.as<float>(), n, dY.as<float>());
        }
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(y, dY.p, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIP_OK(hipStreamSynchronize(m->stream));
    });
}

}  // extern "C"
