// Batched and served reconstruction jobs (lam_server_* in include/lamino.h).
//
// A server owns `slots` worker threads on one device.  Each worker has its
// own context (stream, FFT/NUFFT plans, scratch) and runs one job at a time
// end to end with mbir_run; jobs on different workers overlap on the device.
// Submitted jobs wait in arrival order and are admitted when a worker is free
// and the job's footprint (mbir_footprint) fits in what is left of the HBM
// budget.  A job's device state (iterate, Bregman auxiliaries, Toeplitz
// kernels) stays resident until it finishes; its arithmetic is the same as
// when it runs alone, since every buffer and reduction is the job's own.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "capi_util.h"
#include "nufft.h"
#include "solvers.h"

namespace lam {

namespace {

using clk = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    std::vector<std::vector<float>> proj;
    std::vector<std::vector<double>> alpha, gamma;
    std::vector<lam_dataset> ds;
    long long rd[3] = {0, 0, 0};
    lam_recon_params pr{};
    float *out = nullptr;
    // file output (lam_server_submit_output): path, LAM_OUT_* formats, transposed result
    std::string out_path;
    int out_formats = 0;
    bool out_transpose = false;
    std::string err;
    bool finished = false;
    lam_job_info info{};
};

// Pinned host buffers for results on their way to files: grow-only and reused, since
// freeing pinned memory synchronises the device and would stall the other workers.
struct Staging {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool busy = false;
};

// A finished job's result in pinned memory, for the writer thread.
struct WriteTask {
    Job *job = nullptr;
    Staging *buf = nullptr;
    long long dims[3] = {0, 0, 0};   // of one component, as written
};

}  // namespace

struct Server {
    int device = 0;
    size_t budget = 0;
    int slots = 1;
    MbirMode mode = MbirMode::Plain;
    clk::time_point t0 = clk::now();

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    int64_t next_id = 1, submitted = 0, completed = 0, failed = 0;
    size_t in_use = 0, peak_in_use = 0;
    int running = 0, ready = 0;
    bool stopping = false;
    std::string start_error;
    std::vector<std::thread> workers;
    // file output: pinned staging pool and the writer thread's queue
    std::vector<std::unique_ptr<Staging>> staging;
    std::deque<WriteTask> writes;
    std::condition_variable cv_write, cv_staging;
    bool writer_stop = false;
    std::thread writer;

    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
    // head of the queue fits the budget (callers hold mu)
    bool admissible() const { return !queue.empty() && in_use + queue.front()->info.footprint_est <= budget; }

    void worker(int slot);
    void write_loop();
    Staging *take_staging(size_t bytes);
    void finish(Job *j, lam_status st, const std::string &err);
};

// A free pinned buffer of at least `bytes` (a new one when every free one is
// smaller; waits while all are in use and the pool holds one per worker).
Staging *Server::take_staging(size_t bytes) {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        Staging *best = nullptr;
        for (auto &b : staging)
            if (!b->busy && b->bytes >= bytes && (!best || b->bytes < best->bytes)) best = b.get();
        if (best) {
            best->busy = true;
            return best;
        }
        size_t busy = 0;
        for (auto &b : staging) busy += b->busy ? 1 : 0;
        if (staging.size() < (size_t)(2 * slots) || busy < staging.size()) {
            // grow: replace a free, too small buffer, or add one
            Staging *slot = nullptr;
            for (auto &b : staging)
                if (!b->busy) { slot = b.get(); break; }
            if (!slot) {
                staging.push_back(std::make_unique<Staging>());
                slot = staging.back().get();
            }
            slot->busy = true;
            void *old = slot->ptr;
            slot->ptr = nullptr;
            slot->bytes = 0;
            lk.unlock();
            std::shared_lock<std::shared_mutex> cap(capture_mutex());
            if (old) (void)hipHostFree(old);
            void *p = nullptr;
            const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
            lk.lock();
            if (e != hipSuccess) {
                slot->busy = false;
                throw Error(LAM_ERR_NOMEM, "pinned staging for the result: " + std::string(hipGetErrorString(e)));
            }
            slot->ptr = p;
            slot->bytes = bytes;
            return slot;
        }
        cv_staging.wait(lk);
    }
}

void Server::finish(Job *j, lam_status st, const std::string &err) {
    {
        std::lock_guard<std::mutex> lk(mu);
        j->info.end_s = now();
        j->info.status = st;
        j->err = err;
        j->finished = true;
        if (st == LAM_OK) ++completed; else ++failed;
    }
    cv_done.notify_all();
}

// Writes finished jobs' files (TIFF and/or VTI) while the workers go on with the
// next jobs, then copies the result to the caller's buffer if one was given.
void Server::write_loop() {
    for (;;) {
        WriteTask t;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_write.wait(lk, [&] { return writer_stop || !writes.empty(); });
            if (writes.empty()) return;
            t = writes.front();
            writes.pop_front();
        }
        Job *j = t.job;
        const float *res = static_cast<const float *>(t.buf->ptr);
        const long long m = t.dims[0] * t.dims[1] * t.dims[2];
        const int64_t d[3] = {t.dims[0], t.dims[1], t.dims[2]};
        lam_status st = LAM_OK;
        std::string err;
        const char *path = j->out_path.c_str();
        if (st == LAM_OK && (j->out_formats & LAM_OUT_TIFF))
            st = lam_tiff_write3(path, res, res + m, res + 2 * m, d);
        if (st == LAM_OK && (j->out_formats & LAM_OUT_TIFF_VECTORS))
            st = lam_tiff_write_vectors(path, res, res + m, res + 2 * m, d);
        if (st == LAM_OK && (j->out_formats & LAM_OUT_VTI))
            st = lam_vti_write_vectors(path, res, res + m, res + 2 * m, d, nullptr);
        if (st != LAM_OK) err = lam_last_error();
        if (st == LAM_OK && j->out) std::memcpy(j->out, res, sizeof(float) * 3 * (size_t)m);
        {
            std::lock_guard<std::mutex> lk(mu);
            t.buf->busy = false;
        }
        cv_staging.notify_all();
        finish(j, st, err);
    }
}

void Server::worker(int slot) {
    Ctx cx;
    cx.device = device;
    {
        std::lock_guard<std::mutex> lk(mu);
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&cx.stream, hipStreamNonBlocking);
        if (e != hipSuccess) start_error = hipGetErrorString(e);
        cx.own_stream = e == hipSuccess;
        ++ready;
    }
    cv_done.notify_all();
    if (!cx.own_stream) return;
    alloc_stream() = cx.stream;   // every buffer of this worker is stream-ordered
    MemTally *tally = mem_tally();
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return admissible() || (stopping && queue.empty()); });
            if (!admissible()) return;   // stopping, nothing left
            j = queue.front();
            queue.pop_front();
            in_use += j->info.footprint_est;
            peak_in_use = std::max(peak_in_use, in_use);
            ++running;
            j->info.start_s = now();
            j->info.slot = slot;
        }
        const size_t base = tally->cur;
        tally->peak = tally->cur;
        lam_status st = LAM_OK;
        std::string err;
        Staging *stage = nullptr;
        try {
            float *dst = j->out;
            if (j->out_formats) {
                stage = take_staging(sizeof(float) * 3 * (size_t)(j->rd[0] * j->rd[1] * j->rd[2]));
                dst = static_cast<float *>(stage->ptr);
            }
            mbir_run(cx, j->ds.data(), (int)j->ds.size(), j->rd, j->pr, dst, &j->info.stats, mode, j->out_transpose);
        } catch (const Error &e) {
            st = e.code;
            err = e.what();
        } catch (const std::bad_alloc &e) {
            st = LAM_ERR_NOMEM;
            err = e.what();
        } catch (const std::exception &e) {
            st = LAM_ERR_ARG;
            err = e.what();
        }
        // drop this job's NUFFT plans and scratch (stream-ordered frees, no device
        // synchronisation) so that the worker holds no device memory between jobs
        // beyond its small FFT twiddle tables
        cx.nufft_plans.clear();
        for (DevBuf &w : cx.work) w.release();
        cx.red.release();
        cx.scal.release();
        (void)hipStreamSynchronize(cx.stream);
        const size_t peak = tally->peak - base;
        const bool to_writer = st == LAM_OK && stage;
        {
            std::lock_guard<std::mutex> lk(mu);
            in_use -= j->info.footprint_est;
            --running;
            j->info.footprint_peak = peak;
            j->proj.clear();
            j->proj.shrink_to_fit();
            if (stage && !to_writer) stage->busy = false;
            if (to_writer) {
                WriteTask t;
                t.job = j;
                t.buf = stage;
                if (j->out_transpose) {
                    t.dims[0] = j->rd[1]; t.dims[1] = j->rd[2]; t.dims[2] = j->rd[0];
                } else {
                    t.dims[0] = j->rd[0]; t.dims[1] = j->rd[1]; t.dims[2] = j->rd[2];
                }
                writes.push_back(t);
            }
        }
        cv_work.notify_all();
        if (to_writer) {
            cv_write.notify_all();
        } else {
            cv_staging.notify_all();
            finish(j, st, err);
        }
    }
}

}  // namespace lam

struct lam_server {
    lam::Server impl;
};

using namespace lam;

extern "C" {

lam_status lam_mbir_footprint(const lam_dataset *datasets, int nds, const int64_t recon_dims[3],
                              const lam_recon_params *params, size_t *bytes) {
    return guard([&] {
        need(datasets, "datasets");
        need(params, "params");
        need(bytes, "bytes");
        long long rd[3];
        dims_of(recon_dims, rd);
        *bytes = mbir_footprint(datasets, nds, rd, *params);
    });
}

lam_status lam_server_create(int device, size_t hbm_budget, int slots, int mode, lam_server **out) {
    return guard([&] {
        need(out, "out");
        LAM_REQUIRE(slots >= 1 && slots <= 16, LAM_ERR_ARG, "slots must be in [1, 16]");
        LAM_REQUIRE(mode >= LAM_SERVE_PLAIN && mode <= LAM_SERVE_GRAPH, LAM_ERR_ARG, "unknown serving mode");
        int count = 0;
        LAM_HIP(hipGetDeviceCount(&count));
        LAM_REQUIRE(device >= 0 && device < count, LAM_ERR_ARG, "invalid device ordinal");
        auto s = std::make_unique<lam_server>();
        Server &S = s->impl;
        S.device = device;
        S.slots = slots;
        S.mode = mode == LAM_SERVE_GRAPH ? MbirMode::Graph : mode == LAM_SERVE_ASYNC ? MbirMode::Async : MbirMode::Plain;
        if (hbm_budget == 0) {
            size_t freeb = 0, total = 0;
            LAM_HIP(hipSetDevice(device));
            LAM_HIP(hipMemGetInfo(&freeb, &total));
            hbm_budget = freeb / 10 * 9;
        }
        S.budget = hbm_budget;
        S.writer = std::thread([&S] { S.write_loop(); });
        for (int k = 0; k < slots; ++k) S.workers.emplace_back([&S, k] { S.worker(k); });
        {
            std::unique_lock<std::mutex> lk(S.mu);
            S.cv_done.wait(lk, [&] { return S.ready == slots; });
        }
        if (!S.start_error.empty()) {
            {
                std::lock_guard<std::mutex> lk(S.mu);
                S.stopping = true;
            }
            S.cv_work.notify_all();
            for (auto &t : S.workers) t.join();
            {
                std::lock_guard<std::mutex> lk(S.mu);
                S.writer_stop = true;
            }
            S.cv_write.notify_all();
            S.writer.join();
            throw Error(LAM_ERR_HIP, "server worker start failed: " + S.start_error);
        }
        *out = s.release();
    });
}

lam_status lam_server_destroy(lam_server *server) {
    return guard([&] {
        if (!server) return;
        Server &S = server->impl;
        {
            std::lock_guard<std::mutex> lk(S.mu);
            S.stopping = true;
        }
        S.cv_work.notify_all();
        for (auto &t : S.workers) t.join();
        {
            std::lock_guard<std::mutex> lk(S.mu);
            S.writer_stop = true;   // the writer drains its queue first
        }
        S.cv_write.notify_all();
        S.writer.join();
        for (auto &b : S.staging)
            if (b->ptr) (void)hipHostFree(b->ptr);
        delete server;
    });
}

namespace {
void submit_job(lam_server *server, const lam_dataset *datasets, int nds, const int64_t recon_dims[3],
                const lam_recon_params *params, float *out3_host, const lam_output *output, int64_t *job_id) {
        need(server, "server");
        need(datasets, "datasets");
        need(params, "params");
        need(job_id, "job_id");
        LAM_REQUIRE(out3_host || output, LAM_ERR_ARG, "a job needs out3_host or an output");
        LAM_REQUIRE(nds >= 1, LAM_ERR_ARG, "MBIR needs at least one dataset");
        Server &S = server->impl;
        auto j = std::make_unique<Job>();
        dims_of(recon_dims, j->rd);
        j->pr = *params;
        j->out = out3_host;
        if (output) {
            need(output->path, "output.path");
            LAM_REQUIRE(output->formats > 0 && (output->formats & ~(LAM_OUT_TIFF | LAM_OUT_VTI | LAM_OUT_TIFF_VECTORS)) == 0,
                        LAM_ERR_ARG, "output formats must be a non-empty set of LAM_OUT_* flags");
            j->out_path = output->path;
            j->out_formats = output->formats;
            j->out_transpose = output->transpose120 != 0;
        }
        for (int g = 0; g < nds; ++g) {
            const lam_dataset &D = datasets[g];
            need(D.proj, "dataset.proj");
            need(D.alpha, "dataset.alpha");
            LAM_REQUIRE(D.nproj >= 1 && D.nrows >= 1 && D.ncols >= 1, LAM_ERR_ARG, "invalid dataset shape");
            j->proj.emplace_back(D.proj, D.proj + D.nproj * D.nrows * D.ncols);
            j->alpha.emplace_back(D.alpha, D.alpha + D.nproj);
            j->gamma.emplace_back(D.gamma ? std::vector<double>(D.gamma, D.gamma + D.nproj) : std::vector<double>());
        }
        for (int g = 0; g < nds; ++g) {
            lam_dataset D = datasets[g];
            D.proj = j->proj[g].data();
            D.alpha = j->alpha[g].data();
            D.gamma = j->gamma[g].empty() ? nullptr : j->gamma[g].data();
            j->ds.push_back(D);
        }
        j->info.footprint_est = mbir_footprint(j->ds.data(), nds, j->rd, j->pr);
        LAM_REQUIRE(j->info.footprint_est <= S.budget, LAM_ERR_BUDGET,
                    "job needs more device memory than the server's whole budget");
        {
            std::lock_guard<std::mutex> lk(S.mu);
            LAM_REQUIRE(!S.stopping, LAM_ERR_STATE, "server is stopping");
            j->id = S.next_id++;
            j->info.id = j->id;
            j->info.submit_s = S.now();
            *job_id = j->id;
            S.queue.push_back(j.get());
            S.jobs.emplace(j->id, std::move(j));
            ++S.submitted;
        }
        S.cv_work.notify_all();
}
}  // namespace

lam_status lam_server_submit(lam_server *server, const lam_dataset *datasets, int nds, const int64_t recon_dims[3],
                             const lam_recon_params *params, float *out3_host, int64_t *job_id) {
    return guard([&] {
        need(out3_host, "out3_host");
        submit_job(server, datasets, nds, recon_dims, params, out3_host, nullptr, job_id);
    });
}

lam_status lam_server_submit_output(lam_server *server, const lam_dataset *datasets, int nds,
                                    const int64_t recon_dims[3], const lam_recon_params *params,
                                    const lam_output *output, float *out3_host, int64_t *job_id) {
    return guard([&] {
        need(output, "output");
        submit_job(server, datasets, nds, recon_dims, params, out3_host, output, job_id);
    });
}

lam_status lam_server_wait(lam_server *server, int64_t job_id, lam_job_info *info) {
    std::string err;
    lam_status st = guard([&] {
        need(server, "server");
        Server &S = server->impl;
        std::unique_lock<std::mutex> lk(S.mu);
        auto it = S.jobs.find(job_id);
        LAM_REQUIRE(it != S.jobs.end(), LAM_ERR_ARG, "unknown job id");
        Job *j = it->second.get();
        S.cv_done.wait(lk, [&] { return j->finished; });
        if (info) *info = j->info;
        const lam_status js = j->info.status;
        const std::string jerr = j->err;
        S.jobs.erase(it);   // a job is waited on once; its record goes with it
        if (js != LAM_OK) throw Error(js, jerr);
    });
    return st;
}

lam_status lam_server_query(lam_server *server, lam_server_stats *stats) {
    return guard([&] {
        need(server, "server");
        need(stats, "stats");
        Server &S = server->impl;
        std::lock_guard<std::mutex> lk(S.mu);
        stats->submitted = S.submitted;
        stats->completed = S.completed;
        stats->failed = S.failed;
        stats->running = S.running;
        stats->queued = (int64_t)S.queue.size();
        stats->budget = S.budget;
        stats->in_use = S.in_use;
        stats->peak_in_use = S.peak_in_use;
        stats->slots = S.slots;
        stats->mode = S.mode == MbirMode::Graph ? LAM_SERVE_GRAPH : S.mode == MbirMode::Async ? LAM_SERVE_ASYNC
                                                                                             : LAM_SERVE_PLAIN;
    });
}

}  // extern "C"
