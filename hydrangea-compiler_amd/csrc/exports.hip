// exports.hip - hyd_export_* wrappers (the drop-in boundary) and the GPU
// hyd_sort_indices of the runtime.
//
// Each wrapper is what the HIP backend emits for one exported proc
// (src/Language/Hydrangea/CodegenC.hs:411-425 for the C form): check the
// argument arrays, take their HBM copies, allocate device-resident results,
// launch the fused kernel on the export stream and, with host sync on, copy
// the result back so arr->data is valid on return.
#include <stdarg.h>
#include <stdio.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "../../include/hyd_exports.h"
#include "hyd_internal.hpp"

namespace hyd {
hipStream_t export_stream();
int host_sync_enabled();
void mark_device_written(hyd_array_t* arr);
int sort_indices_i64(const int64_t* keys, int64_t n, int64_t* perm, hipStream_t s, int64_t* ws,
                     int* passes_out);
int64_t sort_workspace_words(int64_t n);
}  // namespace hyd

using namespace hyd;

namespace {

[[noreturn]] void die(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "hydrangea: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    exit(1);
}

void check(int rc, const char* what) {
    if (rc != HYD_OK) die("%s: %s", what, hydhip_last_error());
}

void expect_1d(hyd_array_t* a, size_t elem, const char* who) {
    if (!a) die("%s: null array", who);
    if (a->shape.ndims != 1 || a->elem_size != elem)
        die("%s: expected a 1-D array of %zu-byte elements", who, elem);
}

void expect_2d_f64(hyd_array_t* a, const char* who) {
    if (!a) die("%s: null array", who);
    if (a->shape.ndims != 2 || a->elem_size != sizeof(double))
        die("%s: expected a 2-D float array", who);
}

hyd_tuple_t shape1(int64_t n) {
    hyd_tuple_t t;
    memset(&t, 0, sizeof t);
    t.ndims = 1;
    t.elems[0] = n;
    return t;
}

template <class T>
T* dev(hyd_array_t* a) {
    return static_cast<T*>(hyd_array_device_data(a));
}

void finish(hyd_array_t* out) {
    mark_device_written(out);
    if (host_sync_enabled()) check(hyd_array_sync_host(out), "sync result");
}

}  // namespace

extern "C" {

hyd_array_t* hyd_export_weighted_histogram_bins(hyd_array_t* data, int64_t bins) {
    expect_1d(data, 8, "weighted_histogram");
    if (bins < 1) die("weighted_histogram: bins must be positive");
    hyd_array_t* out = hyd_array_alloc_device(shape1(bins), 8);
    check(hydhip_weighted_histogram_i64(dev<int64_t>(data), hyd_array_length(data), bins,
                                        dev<int64_t>(out), HYD_SCATTER_AUTO, export_stream()),
          "weighted_histogram");
    finish(out);
    return out;
}

hyd_array_t* hyd_export_weighted_histogram(hyd_array_t* data) {
    return hyd_export_weighted_histogram_bins(data, 64);
}

double hyd_export_dot(hyd_array_t* xs, hyd_array_t* ys) {
    expect_1d(xs, 8, "dot");
    expect_1d(ys, 8, "dot");
    // zipwith iterates over the common prefix (shape of the first argument)
    int64_t n = std::min(hyd_array_length(xs), hyd_array_length(ys));
    hipStream_t s = export_stream();
    double* d_out = static_cast<double*>(scalar_slot());
    if (!d_out) die("dot: %s", hydhip_last_error());
    check(hydhip_dot_f64(dev<double>(xs), dev<double>(ys), n, 0.0, d_out, s), "dot");
    double r = 0.0;
    if (hipMemcpyAsync(&r, d_out, sizeof r, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        die("dot: result copy failed");
    return r;
}

hyd_array_t* hyd_export_map_chain(hyd_array_t* xs, hyd_array_t* ys) {
    expect_1d(xs, 8, "map_chain");
    expect_1d(ys, 8, "map_chain");
    int64_t n = std::min(hyd_array_length(xs), hyd_array_length(ys));
    hyd_array_t* out = hyd_array_alloc_device(shape1(n), 8);
    check(hydhip_map_chain_f64(dev<double>(xs), dev<double>(ys), n, dev<double>(out), export_stream()),
          "map_chain");
    finish(out);
    return out;
}

hyd_array_t* hyd_export_heat_step(hyd_array_t* u) {
    expect_2d_f64(u, "heat_step");
    hyd_array_t* out = hyd_array_alloc_device(u->shape, 8);
    check(hydhip_stencil2d_clamp_f64(dev<double>(u), u->shape.elems[0], u->shape.elems[1],
                                     HYD_STENCIL_HEAT, 0.1, dev<double>(out), export_stream()),
          "heat_step");
    finish(out);
    return out;
}

hyd_array_t* hyd_export_laplacian(hyd_array_t* input) {
    expect_2d_f64(input, "laplacian");
    hyd_array_t* out = hyd_array_alloc_device(input->shape, 8);
    check(hydhip_stencil2d_clamp_f64(dev<double>(input), input->shape.elems[0],
                                     input->shape.elems[1], HYD_STENCIL_LAPLACIAN, 0.0,
                                     dev<double>(out), export_stream()),
          "laplacian");
    finish(out);
    return out;
}

hyd_array_t* hyd_export_jacobi_heat(hyd_array_t* u, int64_t n) {
    expect_2d_f64(u, "jacobi_heat");
    if (n < 0) die("jacobi_heat: negative iteration count");
    const int64_t h = u->shape.elems[0], w = u->shape.elems[1];
    hyd_array_t* a = hyd_array_alloc_device(u->shape, 8);
    hyd_array_t* b = n > 1 ? hyd_array_alloc_device(u->shape, 8) : nullptr;
    double* result = nullptr;
    check(hydhip_stencil2d_iterate_f64(dev<double>(u), h, w, HYD_STENCIL_HEAT, 0.1, n, dev<double>(a),
                                       b ? dev<double>(b) : nullptr, /*use_graph=*/1, &result,
                                       export_stream()),
          "jacobi_heat");
    hyd_array_t* out = a;
    if (b && result == dev<double>(b)) {
        out = b;
        b = a;
    }
    if (b) hyd_array_free(b);
    finish(out);
    return out;
}

hyd_record_col_idx_ai_ncols_i_nnz_i_nrows_i_row_ptr_ai_vals_ai_t hyd_export_coo_to_csr(
    int64_t nrows, int64_t ncols, int64_t nnz, hyd_array_t* rows, hyd_array_t* cols,
    hyd_array_t* vals) {
    expect_1d(rows, 8, "coo_to_csr");
    expect_1d(cols, 8, "coo_to_csr");
    expect_1d(vals, 8, "coo_to_csr");
    if (nrows < 0 || ncols < 0 || nnz < 0) die("coo_to_csr: negative size");
    if (nnz > hyd_array_length(rows) || nnz > hyd_array_length(cols) || nnz > hyd_array_length(vals))
        die("coo_to_csr: nnz exceeds the triplet arrays");
    hyd_record_col_idx_ai_ncols_i_nnz_i_nrows_i_row_ptr_ai_vals_ai_t r;
    r.row_ptr = hyd_array_alloc_device(shape1(nrows + 1), 8);
    r.col_idx = hyd_array_alloc_device(shape1(nnz), 8);
    r.vals = hyd_array_alloc_device(shape1(nnz), 8);
    int64_t merged = 0;
    check(hydhip_coo_to_csr_i64(nrows, ncols, nnz, dev<int64_t>(rows), dev<int64_t>(cols),
                                dev<int64_t>(vals), dev<int64_t>(r.row_ptr), dev<int64_t>(r.col_idx),
                                dev<int64_t>(r.vals), &merged, export_stream()),
          "coo_to_csr");
    r.ncols = ncols;
    r.nrows = nrows;
    r.nnz = merged;
    finish(r.row_ptr);
    finish(r.col_idx);
    finish(r.vals);
    return r;
}

hyd_array_t* hyd_sort_indices(hyd_array_t* keys) {
    if (keys->shape.ndims != 1 || keys->elem_size != sizeof(int64_t)) {
        fprintf(stderr, "hyd_sort_indices: expected a 1D int64 array\n");
        exit(1);
    }
    int64_t n = hyd_array_length(keys);
    hyd_array_t* out = hyd_array_alloc_device(keys->shape, 8);
    if (n > 0) {
        hipStream_t s = export_stream();
        int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)sort_workspace_words(n) * 8));
        if (!ws) {
            fprintf(stderr, "hyd_sort_indices: allocation failed\n");
            exit(1);
        }
        check(sort_indices_i64(dev<int64_t>(keys), n, dev<int64_t>(out), s, ws, nullptr),
              "hyd_sort_indices");
    }
    finish(out);
    return out;
}

// The benchmark scaffold the C backend emits in --bench mode
// (src/Language/Hydrangea/CodegenC.hs:515-580): `warmup` untimed calls, then
// `iters` calls each timed with CLOCK_MONOTONIC, then
// "benchmark[<name>]: min=<ms> ms  mean=<ms> ms" on stderr.  Here the body is a
// host callback around hyd_export_* calls, and each timed call also waits for
// the GPU (the library stream and the device) so that kernels left in HBM
// without host sync are inside the measurement.
int64_t hyd_hip_bench(const char* name, void (*body)(void*), void* arg, int64_t warmup, int64_t iters,
                      double* min_ms, double* mean_ms) {
    if (!body || iters <= 0 || warmup < 0) {
        fprintf(stderr, "hyd_hip_bench: bad arguments\n");
        exit(1);
    }
    // a kernel that failed asynchronously surfaces at the wait: the drop-in's
    // error behaviour (message, exit status 1), not a timing of broken work
    auto wait_gpu = [&]() {
        const hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            fprintf(stderr, "hydrangea: hyd_hip_bench(%s): %s\n", name ? name : "", hipGetErrorString(e));
            exit(1);
        }
    };
    for (int64_t w = 0; w < warmup; ++w) body(arg);
    wait_gpu();
    double mn = 0.0, sum = 0.0;
    for (int64_t i = 0; i < iters; ++i) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        body(arg);
        wait_gpu();
        clock_gettime(CLOCK_MONOTONIC, &t1);
        const double ns = (double)(t1.tv_sec - t0.tv_sec) * 1.0e9 + (double)(t1.tv_nsec - t0.tv_nsec);
        if (i == 0 || ns < mn) mn = ns;
        sum += ns;
    }
    fprintf(stderr, "benchmark[%s]: min=%.3f ms  mean=%.3f ms\n", name ? name : "", mn / 1.0e6,
            sum / (double)iters / 1.0e6);
    if (min_ms) *min_ms = mn / 1.0e6;
    if (mean_ms) *mean_ms = sum / (double)iters / 1.0e6;
    return 0;
}

}  // extern "C"
