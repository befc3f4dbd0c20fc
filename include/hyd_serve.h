/* hyd_serve.h - continuous-batching request server over the fused kernels.
 *
 * The reference has no serving layer (its exported kernels are plain,
 * non-reentrant C calls: src/Language/Hydrangea/CodegenC.hs:812-862); this is
 * the MI355X host runtime around the same kernels:
 *   - requests name device-resident inputs (arrays stay in HBM across calls);
 *   - admission by HBM budget: outputs and scratch are carved from one arena
 *     of `hbm_budget_bytes`, so no hipMalloc/hipFree runs per request;
 *     a request that can never fit is rejected at submission;
 *   - continuous batching: requests that arrive while others run wait in a
 *     ready list; when a worker frees up, every ready histogram (or dot)
 *     request becomes one batched launch; stencils and COO->CSR builds run
 *     as their own jobs; workers own one stream each, so jobs overlap;
 *   - results stay in the arena until hyd_server_release.
 * Status codes are those of hyd_hip.h; nothing here calls exit().
 */
#ifndef HYD_SERVE_H
#define HYD_SERVE_H

#include <stdint.h>

#include "hyd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HYD_REQ_STENCIL 0 /* Jacobi heat: in0 = u[h,w] fp64, iters steps -> out0 [h,w] */
#define HYD_REQ_HIST 1    /* weighted histogram: in0 = data[n] int64, bins -> out0 [bins] */
#define HYD_REQ_DOT 2     /* dot: in0 = xs[n], in1 = ys[n] fp64 -> out0 [1] */
#define HYD_REQ_COO 3     /* COO->CSR: in0/in1/in2 = rows/cols/vals [n = nnz] int64,
                             nrows, ncols -> out0 row_ptr[nrows+1], out1 col_idx[n],
                             out2 vals[n]; count = merged nnz */

#define HYD_SERVE_NO_BATCH 1 /* one launch per request (for comparisons) */
#define HYD_SERVE_FAKE 2     /* no GPU: modelled kernel times, address-only arena */

typedef struct {
    int32_t kind;
    int32_t pad;
    int64_t n;
    int64_t h, w, iters;
    int64_t bins;
    int64_t nrows, ncols;
    const void* in0;
    const void* in1;
    const void* in2;
    double deadline_ms; /* > 0: fail with HYD_E_TIMEOUT if not admitted in time */
} hyd_request_t;

typedef struct {
    int32_t status; /* HYD_OK or an error code (final when done == 1) */
    int32_t done;
    void* out0;
    void* out1;
    void* out2;
    int64_t count;
    double submit_ms, admit_ms, start_ms, done_ms; /* server clock */
    int64_t bytes;     /* arena bytes held for the outputs */
    int64_t batch_id;  /* requests launched together share it */
} hyd_result_t;

typedef struct {
    int64_t budget, in_use, peak_in_use;
    int64_t submitted, admitted, completed, failed, rejected, timed_out;
    int64_t queued, running;
    int64_t jobs, batched_jobs, batched_requests;
} hyd_server_stats_t;

typedef struct hyd_server hyd_server_t;

hyd_server_t* hyd_server_create(int64_t hbm_budget_bytes, int workers, int flags);
void hyd_server_destroy(hyd_server_t* s);
/* HYD_E_REJECTED: larger than the whole budget; HYD_E_INVALID: malformed. */
int hyd_server_submit(hyd_server_t* s, const hyd_request_t* req, int64_t* id);
/* HYD_OK when finished (out filled), HYD_E_BUSY while pending. */
int hyd_server_poll(hyd_server_t* s, int64_t id, hyd_result_t* out);
int hyd_server_wait(hyd_server_t* s, int64_t id, double timeout_ms, hyd_result_t* out);
int hyd_server_wait_all(hyd_server_t* s, double timeout_ms);
/* Next finished request in completion order (each reported once):
 * HYD_OK with *id and *out filled, HYD_E_BUSY when none finished in time. */
int hyd_server_next_done(hyd_server_t* s, double timeout_ms, int64_t* id, hyd_result_t* out);
/* Returns the request's outputs to the arena and forgets it. */
int hyd_server_release(hyd_server_t* s, int64_t id);
int hyd_server_stats(hyd_server_t* s, hyd_server_stats_t* out);
/* The next `njobs` jobs fail as if a launch had failed (failure-path tests). */
int hyd_server_inject_failures(hyd_server_t* s, int64_t njobs);
/* Arena bytes a request needs (outputs + scratch). */
int64_t hyd_request_bytes(const hyd_request_t* req);

/* The per-call path the server is compared with: hipMalloc the outputs, run
 * the request's kernels with plain launches, synchronise.  Free with
 * hydhip_result_free_plain.  On any failure nothing stays allocated and
 * *out is left zeroed. */
int hydhip_run_request_plain(const hyd_request_t* req, hyd_result_t* out);
int hydhip_result_free_plain(hyd_result_t* res);
/* Host-only test hook: runs the plain path's allocation sequence for `req`
 * with a counting host allocator whose call number `fail_at` fails (0-based,
 * negative: none); *live_after = allocations still held when it returns. */
int hydhip_test_plain_alloc(const hyd_request_t* req, int fail_at, int64_t* live_after);

#ifdef __cplusplus
}
#endif

#endif /* HYD_SERVE_H */
