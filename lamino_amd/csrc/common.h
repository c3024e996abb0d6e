// Shared definitions for the gfx950 laminography library: error handling,
// the context (device, stream, cached plans, scratch), and small device
// helpers.  Everything here is internal; the public surface is
// include/lamino.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "lamino.h"

namespace lam {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define LAM_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw ::lam::Error(LAM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define LAM_REQUIRE(cond, code, msg)                   \
    do {                                               \
        if (!(cond)) throw ::lam::Error((code), (msg)); \
    } while (0)

// ---------------------------------------------------------------- device buffer
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t b) {
        release();
        if (b) LAM_HIP(hipMalloc(&ptr, b));
        bytes = b;
    }
    // grow-only
    void ensure(size_t b) { if (b > bytes) alloc(b); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct FftPlan;     // fft.hip
struct NufftPlan;   // nufft.hip
struct Ctx;

// HIP-event timing of named kernels on the context stream (off by default).
struct KernelTimer {
    bool on = false;
    std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> ev;
    std::vector<hipEvent_t> pool;
    hipEvent_t take() {
        hipEvent_t e;
        if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
        else LAM_HIP(hipEventCreate(&e));
        return e;
    }
    void clear() {
        for (auto &kv : ev)
            for (auto &pr : kv.second) { pool.push_back(pr.first); pool.push_back(pr.second); }
        ev.clear();
    }
    ~KernelTimer() {
        clear();
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------- context
struct Ctx {
    KernelTimer timer;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::map<std::pair<int, int>, std::unique_ptr<FftPlan>> fft_plans;   // (N, sign)
    std::map<std::tuple<long long, long long, long long, int>, std::unique_ptr<NufftPlan>> nufft_plans;
    // general scratch used by a single call at a time
    DevBuf red;        // FP64 reduction partials
    DevBuf scal;       // small device scalar block
    DevBuf work[4];    // named workspaces (grow-only)
    std::vector<double> host_red;
    ~Ctx();
};

FftPlan &get_fft_plan(Ctx &cx, int n, int sign);

// Records start/stop events around the launches made in its scope when the
// context's timer is on.
struct TimedScope {
    Ctx &cx;
    const char *name;
    hipEvent_t stop = nullptr;
    TimedScope(Ctx &c, const char *n) : cx(c), name(n) {
        if (!cx.timer.on) return;
        hipEvent_t start = cx.timer.take();
        stop = cx.timer.take();
        LAM_HIP(hipEventRecord(start, cx.stream));
        cx.timer.ev[name].push_back({start, stop});
    }
    ~TimedScope() {
        if (stop) (void)hipEventRecord(stop, cx.stream);
    }
};

// Number of blocks for a 1-D grid-stride launch over n items.
inline unsigned grid_for(size_t n, unsigned block = 256, unsigned cap = 8192) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

inline void check_launch() { LAM_HIP(hipGetLastError()); }

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// FP64 block reduction of one value per thread (blockDim.x == 256); result in thread 0.
__device__ __forceinline__ double block_sum_256(double v) {
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

}  // namespace lam

// opaque handles of the C ABI
struct lam_ctx {
    lam::Ctx impl;
};
