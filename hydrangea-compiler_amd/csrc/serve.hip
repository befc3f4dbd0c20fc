// serve.hip - continuous-batching request server over the fused kernels
// (include/hyd_serve.h).
//
// Threads: the client thread submits (and admits what fits right away), one
// scheduler thread re-admits and expires deadlines, and `workers` worker
// threads each own one HIP stream (default priority).  A request goes
//   pending --(arena has room: outputs + scratch carved)--> ready
//   ready   --(a worker is free; all ready HIST or DOT requests become one
//              batched launch, STENCIL/COO run alone)--> running --> done
// Outputs stay in the arena until hyd_server_release; scratch is returned
// when the job ends.  Memory never comes from hipMalloc after creation.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/hyd_serve.h"
#include "hyd_internal.hpp"

namespace hyd {
int batched_histogram_ptrs(const int64_t* const* src, const int64_t* len, const int64_t* bins,
                           int64_t* const* dst, int64_t count, int policy, hipStream_t s,
                           int64_t* out_tiles);
int batched_dot(const double* const* xs, const double* const* ys, const int64_t* lens, int64_t count,
                double init, double* const* outs, hipStream_t s);
int64_t coo_to_csr_workspace_words(int64_t nnz, int64_t nrows);
int coo_to_csr_ws(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* rows, const int64_t* cols,
                  const int64_t* vals, int64_t* row_ptr, int64_t* col_idx, int64_t* out_vals,
                  int64_t* nnz_out, hipStream_t s, int64_t* ws);
}  // namespace hyd

namespace {
using namespace hyd;
using Clock = std::chrono::steady_clock;

constexpr int64_t kGranule = 256;
constexpr size_t kMaxBatch = 1024;   // requests per batched launch
constexpr double kHeadWaitMs = 50.0; // a blocked head request stops others passing it after this
constexpr int kMaxSkip = 64;         // requests admitted past a blocked head per round

int64_t round_up(int64_t b) { return (std::max<int64_t>(b, 1) + kGranule - 1) / kGranule * kGranule; }

// First-fit allocator over [0, size) with a coalescing free list.
class Arena {
  public:
    void init(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
        in_use_ = peak_ = 0;
    }
    int64_t alloc(int64_t bytes) {
        bytes = round_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            int64_t off = it->first, rem = it->second - bytes;
            free_.erase(it);
            if (rem) free_[off + bytes] = rem;
            in_use_ += bytes;
            peak_ = std::max(peak_, in_use_);
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = round_up(bytes);
        in_use_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t in_use() const { return in_use_; }
    int64_t peak() const { return peak_; }

  private:
    std::map<int64_t, int64_t> free_;
    int64_t in_use_ = 0, peak_ = 0;
};

struct Alloc {
    int64_t off, bytes;
    bool keep;  // an output (kept until release) or scratch (freed at the end of the job)
};

enum State { kPending, kReady, kRunning, kDone };

struct Req {
    int64_t id;
    hyd_request_t r;
    hyd_result_t res;
    State state;
    std::vector<Alloc> allocs;
};

// Arena layout of one request: outputs first (kept), then scratch.
std::vector<Alloc> layout(const hyd_request_t& r) {
    std::vector<Alloc> a;
    switch (r.kind) {
        case HYD_REQ_STENCIL: {
            int64_t b = r.h * r.w * 8;
            a.push_back({0, b, true});
            if (r.iters > 1) a.push_back({0, b, false});  // ping-pong partner
            break;
        }
        case HYD_REQ_HIST: a.push_back({0, r.bins * 8, true}); break;
        case HYD_REQ_DOT: a.push_back({0, 8, true}); break;
        case HYD_REQ_COO:
            a.push_back({0, (r.nrows + 1) * 8, true});
            a.push_back({0, r.n * 8, true});
            a.push_back({0, r.n * 8, true});
            a.push_back({0, coo_to_csr_workspace_words(r.n, r.nrows) * 8, false});
            break;
    }
    return a;
}

int validate(const hyd_request_t& r) {
    switch (r.kind) {
        case HYD_REQ_STENCIL:
            if (r.h <= 0 || r.w <= 0 || r.iters < 0 || !r.in0) return HYD_E_INVALID;
            return HYD_OK;
        case HYD_REQ_HIST:
            if (r.n < 0 || r.bins < 1 || r.bins > ((int64_t)1 << 31) || (r.n > 0 && !r.in0)) return HYD_E_INVALID;
            return HYD_OK;
        case HYD_REQ_DOT:
            if (r.n < 0 || (r.n > 0 && (!r.in0 || !r.in1))) return HYD_E_INVALID;
            return HYD_OK;
        case HYD_REQ_COO:
            if (r.n < 0 || r.nrows < 0 || r.ncols < 0 || (r.n > 0 && (!r.in0 || !r.in1 || !r.in2)))
                return HYD_E_INVALID;
            return HYD_OK;
    }
    return HYD_E_INVALID;
}

// ---- the plain per-call path's buffers ----------------------------------
// Byte sizes of the device buffers one plain call holds, in output order
// (the stencil's second entry is the ping-pong partner, 0 when one step or
// none needs no partner).  Returns how many there are.
int plain_buffer_sizes(const hyd_request_t& r, int64_t sizes[3]) {
    switch (r.kind) {
        case HYD_REQ_STENCIL:
            sizes[0] = r.h * r.w * 8;
            sizes[1] = r.iters > 1 ? r.h * r.w * 8 : 0;
            return 2;
        case HYD_REQ_HIST: sizes[0] = r.bins * 8; return 1;
        case HYD_REQ_DOT: sizes[0] = 8; return 1;
        default:
            sizes[0] = (r.nrows + 1) * 8;
            sizes[1] = r.n * 8;
            sizes[2] = r.n * 8;
            return 3;
    }
}

using AllocFn = int (*)(void**, int64_t);
using FreeFn = void (*)(void*);

// Allocates every nonzero size in order; on a failure frees what it had
// allocated, leaves every pointer null and returns the failure's status.
int plain_alloc_all(const int64_t* sizes, int n, void** ptrs, AllocFn alloc, FreeFn release) {
    for (int k = 0; k < n; ++k) ptrs[k] = nullptr;
    for (int k = 0; k < n; ++k) {
        if (sizes[k] <= 0) continue;
        const int rc = alloc(&ptrs[k], sizes[k]);
        if (rc != HYD_OK) {
            for (int j = 0; j < k; ++j) {
                release(ptrs[j]);
                ptrs[j] = nullptr;
            }
            ptrs[k] = nullptr;
            return rc;
        }
    }
    return HYD_OK;
}

int device_alloc(void** p, int64_t bytes) {
    if (hipMalloc(p, (size_t)std::max<int64_t>(bytes, 8)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return set_error(HYD_E_OOM, "hipMalloc(%lld) failed", (long long)bytes);
    }
    return HYD_OK;
}

void device_free(void* p) {
    if (p) (void)hipFree(p);
}

// The host-only allocator of hydhip_test_plain_alloc.
int g_test_fail_at = -1, g_test_calls = 0;
int64_t g_test_live = 0;

int test_alloc(void** p, int64_t bytes) {
    if (g_test_calls++ == g_test_fail_at) {
        *p = nullptr;
        return set_error(HYD_E_OOM, "injected allocation failure (%lld bytes)", (long long)bytes);
    }
    *p = malloc((size_t)std::max<int64_t>(bytes, 8));
    if (!*p) return set_error(HYD_E_OOM, "malloc(%lld) failed", (long long)bytes);
    ++g_test_live;
    return HYD_OK;
}

void test_free(void* p) {
    if (!p) return;
    free(p);
    --g_test_live;
}

}  // namespace

struct hyd_server {
    int flags = 0;
    bool fake = false;
    int64_t budget = 0;
    unsigned char* base = nullptr;  // arena base (HBM, or a fake address)
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_sched, cv_work, cv_done;
    std::deque<Req*> pending, ready;
    std::deque<int64_t> finished;  // completion order, drained by hyd_server_next_done
    std::unordered_map<int64_t, Req*> reqs;
    std::vector<std::thread> workers;
    std::thread sched;
    std::vector<hipStream_t> streams;
    bool stop = false;
    int64_t next_id = 1, next_batch = 1, inject = 0;
    hyd_server_stats_t st{};
    Clock::time_point epoch = Clock::now();
    size_t max_ready = 4096;

    double now() const {
        return std::chrono::duration<double, std::milli>(Clock::now() - epoch).count();
    }
    void* addr(int64_t off) const { return base + off; }
};

namespace {

void finish_locked(hyd_server* S, Req* q, int status) {
    // scratch always goes back; outputs too when the request failed
    for (const Alloc& a : q->allocs)
        if (!a.keep || status != HYD_OK) S->arena.release(a.off, a.bytes);
    if (status != HYD_OK) {
        q->allocs.clear();
        q->res.out0 = q->res.out1 = q->res.out2 = nullptr;
        q->res.bytes = 0;
        ++S->st.failed;
    } else {
        std::vector<Alloc> kept;
        int64_t held = 0;
        for (const Alloc& a : q->allocs)
            if (a.keep) {
                kept.push_back(a);
                held += round_up(a.bytes);
            }
        q->allocs.swap(kept);
        q->res.bytes = held;
        ++S->st.completed;
    }
    q->res.status = status;
    q->res.done = 1;
    q->res.done_ms = S->now();
    q->state = kDone;
    S->finished.push_back(q->id);
}

bool try_allocate(hyd_server* S, Req* q) {
    std::vector<Alloc> want = layout(q->r);
    size_t i = 0;
    for (; i < want.size(); ++i) {
        int64_t off = S->arena.alloc(want[i].bytes);
        if (off < 0) break;
        want[i].off = off;
    }
    if (i < want.size()) {
        for (size_t j = 0; j < i; ++j) S->arena.release(want[j].off, want[j].bytes);
        return false;
    }
    q->allocs = want;
    return true;
}

// Expire deadlines and admit what fits (called with the lock held).
void admit_locked(hyd_server* S) {
    const double t = S->now();
    for (auto it = S->pending.begin(); it != S->pending.end();) {
        Req* q = *it;
        if (q->r.deadline_ms > 0 && t - q->res.submit_ms > q->r.deadline_ms) {
            finish_locked(S, q, HYD_E_TIMEOUT);
            ++S->st.timed_out;
            --S->st.failed;  // counted as timed out, not failed
            it = S->pending.erase(it);
        } else {
            ++it;
        }
    }
    int skipped = 0;
    for (auto it = S->pending.begin(); it != S->pending.end();) {
        if (S->ready.size() >= S->max_ready) break;
        Req* q = *it;
        if (!try_allocate(S, q)) {
            // the head may be passed by smaller requests, but not forever
            if (it == S->pending.begin() && t - q->res.submit_ms > kHeadWaitMs) break;
            if (++skipped > kMaxSkip) break;
            ++it;
            continue;
        }
        q->state = kReady;
        q->res.admit_ms = t;
        ++S->st.admitted;
        S->ready.push_back(q);
        it = S->pending.erase(it);
    }
    if (!S->ready.empty()) S->cv_work.notify_all();
}

// Runs one job on `s` and waits for it.  Called without the lock: the
// requests of a running job are touched only by this worker.
int run_job(hyd_server* S, std::vector<Req*>& job, hipStream_t s) {
    const int kind = job[0]->r.kind;
    if (S->fake) {
        // modelled device time: compulsory bytes at 4 TB/s + 5 us per launch
        double bytes = 0;
        for (Req* q : job) {
            const hyd_request_t& r = q->r;
            if (kind == HYD_REQ_STENCIL) bytes += 16.0 * r.h * r.w * std::max<int64_t>(r.iters, 1);
            if (kind == HYD_REQ_HIST) bytes += 8.0 * r.n + 8.0 * r.bins;
            if (kind == HYD_REQ_DOT) bytes += 16.0 * r.n;
            if (kind == HYD_REQ_COO) bytes += 240.0 * r.n;
        }
        double us = bytes / 4e6 + 5.0;
        std::this_thread::sleep_for(std::chrono::microseconds((int64_t)us));
        for (Req* q : job) {
            q->res.out0 = S->addr(q->allocs[0].off);
            if (kind == HYD_REQ_COO) {
                q->res.out1 = S->addr(q->allocs[1].off);
                q->res.out2 = S->addr(q->allocs[2].off);
                q->res.count = q->r.n;
            }
        }
        return HYD_OK;
    }
    int rc = HYD_OK;
    if (kind == HYD_REQ_HIST) {
        std::vector<const int64_t*> src;
        std::vector<int64_t> len, bins;
        std::vector<int64_t*> dst;
        for (Req* q : job) {
            src.push_back(static_cast<const int64_t*>(q->r.in0));
            len.push_back(q->r.n);
            bins.push_back(q->r.bins);
            dst.push_back(static_cast<int64_t*>(S->addr(q->allocs[0].off)));
            q->res.out0 = dst.back();
        }
        rc = batched_histogram_ptrs(src.data(), len.data(), bins.data(), dst.data(), (int64_t)job.size(),
                                    0, s, nullptr);
    } else if (kind == HYD_REQ_DOT) {
        std::vector<const double*> xs, ys;
        std::vector<int64_t> lens;
        std::vector<double*> outs;
        for (Req* q : job) {
            xs.push_back(static_cast<const double*>(q->r.in0));
            ys.push_back(static_cast<const double*>(q->r.in1));
            lens.push_back(q->r.n);
            outs.push_back(static_cast<double*>(S->addr(q->allocs[0].off)));
            q->res.out0 = outs.back();
        }
        rc = batched_dot(xs.data(), ys.data(), lens.data(), (int64_t)job.size(), 0.0, outs.data(), s);
    } else if (kind == HYD_REQ_STENCIL) {
        Req* q = job[0];
        double* b0 = static_cast<double*>(S->addr(q->allocs[0].off));
        double* b1 = q->allocs.size() > 1 ? static_cast<double*>(S->addr(q->allocs[1].off)) : nullptr;
        double* result = nullptr;
        rc = hydhip_stencil2d_iterate_f64(static_cast<const double*>(q->r.in0), q->r.h, q->r.w,
                                          HYD_STENCIL_HEAT, 0.1, q->r.iters, b0, b1, 0, &result, s);
        if (rc == HYD_OK && result == b1) {
            // the result landed in the partner buffer: keep that one instead
            q->allocs[0].keep = false;
            q->allocs[1].keep = true;
        }
        q->res.out0 = result;
    } else if (kind == HYD_REQ_COO) {
        Req* q = job[0];
        int64_t cnt = 0;
        auto* rp = static_cast<int64_t*>(S->addr(q->allocs[0].off));
        auto* ci = static_cast<int64_t*>(S->addr(q->allocs[1].off));
        auto* cv = static_cast<int64_t*>(S->addr(q->allocs[2].off));
        auto* ws = static_cast<int64_t*>(S->addr(q->allocs[3].off));
        rc = coo_to_csr_ws(q->r.nrows, q->r.ncols, q->r.n, static_cast<const int64_t*>(q->r.in0),
                           static_cast<const int64_t*>(q->r.in1), static_cast<const int64_t*>(q->r.in2),
                           rp, ci, cv, &cnt, s, ws);
        q->res.out0 = rp;
        q->res.out1 = ci;
        q->res.out2 = cv;
        q->res.count = cnt;
    }
    hipError_t e = hipStreamSynchronize(s);
    if (rc == HYD_OK && e != hipSuccess) rc = set_error(HYD_E_HIP, "job failed: %s", hipGetErrorString(e));
    return rc;
}

void worker_loop(hyd_server* S, int w) {
    hipStream_t s = S->fake ? nullptr : S->streams[(size_t)w];
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_work.wait(lk, [&] { return S->stop || !S->ready.empty(); });
        if (S->stop) break;
        // the oldest ready request decides the job; HIST/DOT take every
        // ready request of their kind with them (continuous batching)
        Req* head = S->ready.front();
        std::vector<Req*> job;
        const int kind = head->r.kind;
        if (!(S->flags & HYD_SERVE_NO_BATCH) && (kind == HYD_REQ_HIST || kind == HYD_REQ_DOT)) {
            for (auto it = S->ready.begin(); it != S->ready.end() && job.size() < kMaxBatch;) {
                if ((*it)->r.kind == kind) {
                    job.push_back(*it);
                    it = S->ready.erase(it);
                } else {
                    ++it;
                }
            }
        } else {
            job.push_back(head);
            S->ready.pop_front();
        }
        const int64_t bid = S->next_batch++;
        const double t = S->now();
        for (Req* q : job) {
            q->state = kRunning;
            q->res.start_ms = t;
            q->res.batch_id = bid;
        }
        S->st.running += (int64_t)job.size();
        ++S->st.jobs;
        if (job.size() > 1) {
            ++S->st.batched_jobs;
            S->st.batched_requests += (int64_t)job.size();
        }
        const bool fail = S->inject > 0;
        if (fail) --S->inject;
        lk.unlock();
        int rc = fail ? HYD_E_HIP : run_job(S, job, s);
        lk.lock();
        for (Req* q : job) finish_locked(S, q, rc);
        S->st.running -= (int64_t)job.size();
        admit_locked(S);  // memory came back
        S->cv_done.notify_all();
    }
}

void sched_loop(hyd_server* S) {
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->stop) {
        S->cv_sched.wait_for(lk, std::chrono::milliseconds(1));
        if (S->stop) break;
        admit_locked(S);
    }
}

}  // namespace

extern "C" {

int64_t hyd_request_bytes(const hyd_request_t* req) {
    if (!req || validate(*req) != HYD_OK) return -1;
    int64_t b = 0;
    for (const Alloc& a : layout(*req)) b += round_up(a.bytes);
    return b;
}

hyd_server_t* hyd_server_create(int64_t hbm_budget_bytes, int workers, int flags) {
    if (hbm_budget_bytes <= 0 || workers <= 0 || workers > 32) {
        set_error(HYD_E_INVALID, "hyd_server_create: bad arguments");
        return nullptr;
    }
    hyd_server* S = new hyd_server();
    S->flags = flags;
    S->fake = (flags & HYD_SERVE_FAKE) != 0;
    S->budget = hbm_budget_bytes;
    if (S->fake) {
        S->base = reinterpret_cast<unsigned char*>((uintptr_t)1 << 40);  // addresses only
    } else {
        if (ensure_init() != HYD_OK) {
            delete S;
            return nullptr;
        }
        void* p = nullptr;
        if (hipMalloc(&p, (size_t)hbm_budget_bytes) != hipSuccess) {
            (void)hipGetLastError();
            set_error(HYD_E_OOM, "hyd_server_create: arena of %lld bytes not available",
                      (long long)hbm_budget_bytes);
            delete S;
            return nullptr;
        }
        S->base = static_cast<unsigned char*>(p);
        for (int w = 0; w < workers; ++w) {
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
                set_error(HYD_E_HIP, "hyd_server_create: stream creation failed");
                for (hipStream_t t : S->streams) (void)hydhip_stream_destroy(t);
                (void)hipFree(p);
                delete S;
                return nullptr;
            }
            // size the stream's plan workspace up front (no growth, hence no
            // hipFree, while serving)
            (void)stream_workspace(s, (size_t)16 << 20);
            S->streams.push_back(s);
        }
    }
    S->arena.init(hbm_budget_bytes);
    S->st.budget = hbm_budget_bytes;
    for (int w = 0; w < workers; ++w) S->workers.emplace_back(worker_loop, S, w);
    S->sched = std::thread(sched_loop, S);
    return S;
}

void hyd_server_destroy(hyd_server_t* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_sched.notify_all();
    for (auto& t : S->workers) t.join();
    S->sched.join();
    for (auto& kv : S->reqs) delete kv.second;
    if (!S->fake) {
        for (hipStream_t s : S->streams) (void)hydhip_stream_destroy(s);
        (void)hydhip_free(S->base);
    }
    delete S;
}

int hyd_server_submit(hyd_server_t* S, const hyd_request_t* req, int64_t* id) {
    if (!S || !req || !id) return set_error(HYD_E_INVALID, "hyd_server_submit: null argument");
    if (validate(*req) != HYD_OK) return set_error(HYD_E_INVALID, "hyd_server_submit: malformed request");
    std::lock_guard<std::mutex> lk(S->mu);
    ++S->st.submitted;
    if (hyd_request_bytes(req) > S->budget) {
        ++S->st.rejected;
        return set_error(HYD_E_REJECTED, "request needs %lld bytes, budget is %lld",
                         (long long)hyd_request_bytes(req), (long long)S->budget);
    }
    Req* q = new Req();
    q->id = S->next_id++;
    q->r = *req;
    memset(&q->res, 0, sizeof q->res);
    q->res.submit_ms = S->now();
    q->state = kPending;
    S->reqs[q->id] = q;
    S->pending.push_back(q);
    *id = q->id;
    admit_locked(S);
    return HYD_OK;
}

int hyd_server_poll(hyd_server_t* S, int64_t id, hyd_result_t* out) {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return set_error(HYD_E_INVALID, "unknown request %lld", (long long)id);
    if (out) *out = it->second->res;
    return it->second->state == kDone ? HYD_OK : HYD_E_BUSY;
}

int hyd_server_wait(hyd_server_t* S, int64_t id, double timeout_ms, hyd_result_t* out) {
    std::unique_lock<std::mutex> lk(S->mu);
    auto deadline = Clock::now() + std::chrono::microseconds((int64_t)(timeout_ms * 1000));
    for (;;) {
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) return set_error(HYD_E_INVALID, "unknown request %lld", (long long)id);
        if (it->second->state == kDone) {
            if (out) *out = it->second->res;
            return HYD_OK;
        }
        if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            auto jt = S->reqs.find(id);
            if (jt != S->reqs.end() && jt->second->state == kDone) continue;
            return HYD_E_BUSY;
        }
    }
}

int hyd_server_wait_all(hyd_server_t* S, double timeout_ms) {
    std::unique_lock<std::mutex> lk(S->mu);
    auto deadline = Clock::now() + std::chrono::microseconds((int64_t)(timeout_ms * 1000));
    for (;;) {
        bool busy = !S->pending.empty() || !S->ready.empty() || S->st.running > 0;
        if (!busy) return HYD_OK;
        if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            busy = !S->pending.empty() || !S->ready.empty() || S->st.running > 0;
            return busy ? HYD_E_BUSY : HYD_OK;
        }
    }
}

int hyd_server_next_done(hyd_server_t* S, double timeout_ms, int64_t* id, hyd_result_t* out) {
    std::unique_lock<std::mutex> lk(S->mu);
    auto deadline = Clock::now() + std::chrono::microseconds((int64_t)(timeout_ms * 1000));
    while (S->finished.empty()) {
        if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && S->finished.empty())
            return HYD_E_BUSY;
    }
    const int64_t rid = S->finished.front();
    S->finished.pop_front();
    if (id) *id = rid;
    auto it = S->reqs.find(rid);
    if (out && it != S->reqs.end()) *out = it->second->res;
    return HYD_OK;
}

int hyd_server_release(hyd_server_t* S, int64_t id) {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return set_error(HYD_E_INVALID, "unknown request %lld", (long long)id);
    Req* q = it->second;
    if (q->state != kDone) return set_error(HYD_E_BUSY, "request %lld still running", (long long)id);
    for (const Alloc& a : q->allocs) S->arena.release(a.off, a.bytes);
    delete q;
    S->reqs.erase(it);
    admit_locked(S);
    return HYD_OK;
}

int hyd_server_stats(hyd_server_t* S, hyd_server_stats_t* out) {
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->in_use = S->arena.in_use();
    out->peak_in_use = S->arena.peak();
    out->queued = (int64_t)(S->pending.size() + S->ready.size());
    return HYD_OK;
}

int hyd_server_inject_failures(hyd_server_t* S, int64_t njobs) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->inject = njobs;
    return HYD_OK;
}

int hydhip_run_request_plain(const hyd_request_t* req, hyd_result_t* out) {
    if (!req || !out) return set_error(HYD_E_INVALID, "hydhip_run_request_plain: null argument");
    memset(out, 0, sizeof *out);
    if (validate(*req) != HYD_OK) return set_error(HYD_E_INVALID, "hydhip_run_request_plain: malformed request");
    HYD_TRY(ensure_init());
    const hyd_request_t& r = *req;
    hipStream_t s = resolve_stream(nullptr);
    int64_t sizes[3];
    void* buf[3] = {nullptr, nullptr, nullptr};
    const int nbuf = plain_buffer_sizes(r, sizes);
    HYD_TRY(plain_alloc_all(sizes, nbuf, buf, device_alloc, device_free));
    int rc = HYD_OK;
    int64_t count = 0;
    double* result = nullptr;
    if (r.kind == HYD_REQ_STENCIL) {
        rc = hydhip_stencil2d_iterate_f64(static_cast<const double*>(r.in0), r.h, r.w, HYD_STENCIL_HEAT,
                                          0.1, r.iters, static_cast<double*>(buf[0]),
                                          static_cast<double*>(buf[1]), 0, &result, s);
    } else if (r.kind == HYD_REQ_HIST) {
        rc = hydhip_weighted_histogram_i64(static_cast<const int64_t*>(r.in0), r.n, r.bins,
                                           static_cast<int64_t*>(buf[0]), HYD_SCATTER_AUTO, s);
    } else if (r.kind == HYD_REQ_DOT) {
        rc = hydhip_dot_f64(static_cast<const double*>(r.in0), static_cast<const double*>(r.in1), r.n, 0.0,
                            static_cast<double*>(buf[0]), s);
    } else {
        rc = hydhip_coo_to_csr_i64(r.nrows, r.ncols, r.n, static_cast<const int64_t*>(r.in0),
                                   static_cast<const int64_t*>(r.in1), static_cast<const int64_t*>(r.in2),
                                   static_cast<int64_t*>(buf[0]), static_cast<int64_t*>(buf[1]),
                                   static_cast<int64_t*>(buf[2]), &count, s);
    }
    // wait for the work even after a failed call: kernels queued before the
    // failure may still be reading or writing the buffers about to be freed
    const hipError_t se = hipStreamSynchronize(s);
    if (rc == HYD_OK && se != hipSuccess)
        rc = set_error(HYD_E_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(se));
    if (rc != HYD_OK) {
        for (int k = 0; k < nbuf; ++k) device_free(buf[k]);
        return rc;
    }
    if (r.kind == HYD_REQ_STENCIL && result == static_cast<double*>(buf[1])) std::swap(buf[0], buf[1]);
    if (r.kind == HYD_REQ_STENCIL) {
        device_free(buf[1]);  // the partner buffer of the ping-pong
        buf[1] = nullptr;
    }
    out->out0 = buf[0];
    out->out1 = buf[1];
    out->out2 = buf[2];
    out->count = count;
    out->status = HYD_OK;
    out->done = 1;
    return HYD_OK;
}

// Host-only test of the plain path's allocation sequence: a counting host
// allocator whose call number `fail_at` (0-based; negative: none) fails.
// Returns the sequence's status; *live_after is the number of allocations
// still held afterwards (the buffers on success, 0 after a failure).
int hydhip_test_plain_alloc(const hyd_request_t* req, int fail_at, int64_t* live_after) {
    if (!req || !live_after) return set_error(HYD_E_INVALID, "hydhip_test_plain_alloc: null argument");
    if (validate(*req) != HYD_OK) return set_error(HYD_E_INVALID, "hydhip_test_plain_alloc: malformed request");
    int64_t sizes[3];
    void* buf[3] = {nullptr, nullptr, nullptr};
    const int nbuf = plain_buffer_sizes(*req, sizes);
    g_test_fail_at = fail_at;
    g_test_calls = 0;
    g_test_live = 0;
    const int rc = plain_alloc_all(sizes, nbuf, buf, test_alloc, test_free);
    *live_after = g_test_live;
    for (int k = 0; k < nbuf; ++k) test_free(buf[k]);
    return rc;
}

int hydhip_result_free_plain(hyd_result_t* res) {
    if (!res) return HYD_OK;
    void* ptrs[3] = {res->out0, res->out1, res->out2};
    for (void* p : ptrs)
        if (p) HYD_TRY_HIP(hipFree(p));
    res->out0 = res->out1 = res->out2 = nullptr;
    return HYD_OK;
}

}  // extern "C"
