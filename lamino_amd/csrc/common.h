// Shared definitions for the gfx950 laminography library: error handling,
// the context (device, stream, cached plans, scratch), and small device
// helpers.  Everything here is internal; the public surface is
// include/lamino.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
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
// Device bytes held by the buffers a host thread allocated: current and peak.
// The job server runs one job at a time per worker thread and reads the peak
// as the job's measured footprint.  (Never freed: buffers may outlive the
// thread-exit order of thread_local objects.)
struct MemTally {
    size_t cur = 0, peak = 0;
};
inline MemTally *mem_tally() {
    thread_local MemTally *t = new MemTally;
    return t;
}

// Graph capture (GraphStep) and device allocation exclude each other across
// threads: hipMalloc / hipFree may synchronise the device, which must not meet
// a stream that another thread is capturing.  Allocating on a thread that is
// itself capturing is a bug and raises instead of waiting on itself.
inline std::shared_mutex &capture_mutex() {
    static std::shared_mutex m;
    return m;
}
inline bool &capturing_here() {
    thread_local bool c = false;
    return c;
}

// Stream for stream-ordered allocation on this host thread (job-server
// workers set their own stream): hipMallocAsync / hipFreeAsync on it never
// synchronise the device, so concurrent jobs do not stall each other.  Unset,
// buffers use hipMalloc / hipFree.
inline hipStream_t &alloc_stream() {
    thread_local hipStream_t s = nullptr;
    return s;
}

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    MemTally *tally = nullptr;
    hipStream_t stream = nullptr;   // the stream it was allocated on, when stream-ordered
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes), tally(o.tally), stream(o.stream) {
        o.ptr = nullptr; o.bytes = 0; o.tally = nullptr; o.stream = nullptr;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr; bytes = o.bytes; tally = o.tally; stream = o.stream;
            o.ptr = nullptr; o.bytes = 0; o.tally = nullptr; o.stream = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t b) {
        release();
        if (b) {
            if (capturing_here()) throw Error(LAM_ERR_STATE, "device allocation inside a graph capture");
            std::shared_lock<std::shared_mutex> lk(capture_mutex());
            stream = alloc_stream();
            if (stream) LAM_HIP(hipMallocAsync(&ptr, b, stream));
            else LAM_HIP(hipMalloc(&ptr, b));
            tally = mem_tally();
            tally->cur += b;
            if (tally->cur > tally->peak) tally->peak = tally->cur;
        }
        bytes = b;
    }
    // grow-only
    void ensure(size_t b) { if (b > bytes) alloc(b); }
    void release() {
        if (ptr) {
            std::shared_lock<std::shared_mutex> lk(capture_mutex());
            if (stream) (void)hipFreeAsync(ptr, stream);
            else (void)hipFree(ptr);
            if (tally) tally->cur -= bytes;
        }
        ptr = nullptr;
        bytes = 0;
        tally = nullptr;
        stream = nullptr;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct FftPlan;     // fft.hip
struct NufftPlan;   // nufft.hip
struct Ctx;

// HIP-event timing of named kernels on the context stream (off by default).
struct KernelTimer {
    bool on = false;
    std::string only;   // when non-empty: record only scopes of this name
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
    // spreader/interpolator tile shape per (fine grid, width, components, detector), from the autotuner
    std::map<std::vector<long long>, std::array<int, 3>> bin_tune;
    // general scratch used by a single call at a time
    DevBuf red;        // FP64 reduction partials
    DevBuf scal;       // small device scalar block
    DevBuf work[4];    // named workspaces (grow-only)
    std::vector<double> host_red;
    ~Ctx();
};

FftPlan &get_fft_plan(Ctx &cx, int n, int sign);

// Host -> device copy ordered on the context stream and complete on return.
// It never touches the legacy stream, so it is safe while another thread of
// the process is capturing a graph.
inline void upload_now(Ctx &cx, void *dst, const void *src, size_t bytes) {
    LAM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
}

// Records start/stop events around the launches made in its scope when the
// context's timer is on.
struct TimedScope {
    Ctx &cx;
    const char *name;
    hipEvent_t stop = nullptr;
    TimedScope(Ctx &c, const char *n) : cx(c), name(n) {
        if (!cx.timer.on || (!cx.timer.only.empty() && cx.timer.only != name)) return;
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

// The hardware deals workgroups to the 8 XCDs (each with its own L2) in turn by
// linear id, so neighbouring work items, which often share cache lines (the
// column tiles of an inner-index FFT pass, the overlapping boxes of adjacent 2-D
// tiles), would each land on a different XCD and each fetch the shared lines from
// HBM.  xcd_order(lin, total) renumbers so that each XCD takes a contiguous run of
// work items: a bijection on [0, total) (the last total % 8 keep their numbers).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_order(unsigned lin, unsigned total) {
    const unsigned full = total / kXcds * kXcds;
    return lin < full ? (lin % kXcds) * (full / kXcds) + lin / kXcds : lin;
}

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
