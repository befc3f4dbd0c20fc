/* hyd_hip_glue.h - the calls that the HIP offload pass emits.
 *
 * The pass (hydrangea-compiler_amd/haskell/Language/Hydrangea/CodegenHIP.hs)
 * replaces each loop of a shape listed below with one runtime call
 * (an RCall, which src/Language/Hydrangea/CodegenC.hs:2340-2344 prints as a
 * plain C call).  The C that CodegenC prints is then compiled with
 * `-include hyd_hip_glue.h` and linked with -lhydhip (HIPBackend.hs).
 *
 * Every helper takes the hyd_array_t* the loop read and wrote, leaves
 * arr->data valid on return as the loop did, and hands the array back to the
 * host: the printed C around the call reads and writes arr->data without
 * telling the library, so the next helper uploads it again rather than trust
 * an HBM copy that may be stale.  Each returns 0 (the call's value
 * lands in the variable the pass names, like the reference's
 * `_discarded = hyd_write_array_csv(...)`, Lowering.hs:1911-1915).  A failure
 * exits with a message, as the reference runtime's errors do.
 */
#ifndef HYD_HIP_GLUE_H
#define HYD_HIP_GLUE_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "hyd_hip_runtime.h"

static inline void hydg_die(const char* who) {
    fprintf(stderr, "%s: %s\n", who, hydhip_last_error());
    exit(1);
}

static inline void hydg_check_len(const char* who, hyd_array_t* a, int64_t n) {
    if (hyd_shape_size(a->shape) < n) {
        fprintf(stderr, "%s: the loop runs over %lld elements, an array holds %lld\n", who, (long long)n,
                (long long)hyd_shape_size(a->shape));
        exit(1);
    }
}

/* for i in [0, n): dst[idx[i]] += vals[i] on int64 (mod 2^64): the atomic or
 * privatized scatter-add loop, CodegenC.hs:1084-1288 (detected by the shape
 * of CodegenC.hs:1576-1611, unguarded, index and value loaded at i).  A route
 * outside dst is dropped where the C loop would write out of bounds. */
static inline int64_t hydg_scatter_add_i64(hyd_array_t* dst, hyd_array_t* idx, hyd_array_t* vals, int64_t n) {
    if (n <= 0) return 0;
    hydg_check_len("hydg_scatter_add_i64", idx, n);
    hydg_check_len("hydg_scatter_add_i64", vals, n);
    const int64_t* ix = (const int64_t*)hyd_array_device_data(idx);
    const int64_t* v = (const int64_t*)hyd_array_device_data(vals);
    int64_t* d = (int64_t*)hyd_array_device_data(dst);
    if (hydhip_scatter_add_i64(ix, v, n, d, hyd_shape_size(dst->shape), 0, NULL) != 0)
        hydg_die("hydg_scatter_add_i64");
    hyd_array_mark_device_written(dst);
    if (hyd_array_sync_host(dst) != 0) hydg_die("hydg_scatter_add_i64");
    hyd_array_mark_host_dirty(dst);
    return 0;
}

/* The same loop on fp64 ((+.)): bit-exact to the serial loop, each bin
 * receiving its values in element order. */
static inline int64_t hydg_scatter_add_f64(hyd_array_t* dst, hyd_array_t* idx, hyd_array_t* vals, int64_t n) {
    if (n <= 0) return 0;
    hydg_check_len("hydg_scatter_add_f64", idx, n);
    hydg_check_len("hydg_scatter_add_f64", vals, n);
    const int64_t* ix = (const int64_t*)hyd_array_device_data(idx);
    const double* v = (const double*)hyd_array_device_data(vals);
    double* d = (double*)hyd_array_device_data(dst);
    if (hydhip_scatter_guarded_add_f64(ix, v, NULL, n, d, hyd_shape_size(dst->shape), NULL) != 0)
        hydg_die("hydg_scatter_add_f64");
    hyd_array_mark_device_written(dst);
    if (hyd_array_sync_host(dst) != 0) hydg_die("hydg_scatter_add_f64");
    hyd_array_mark_host_dirty(dst);
    return 0;
}

/* for i in [0, n): x = data[i]; dst[x % nbins] += x * x on int64 (mod 2^64):
 * the fused weighted histogram `scatter (+) hist (map (x % nbins) data)
 * (map (x * x) data)`, CodegenC.hs:1119-1288.  The kernel counts into a
 * zeroed device histogram of nbins bins, which is added into dst on the
 * host; a route outside [0, nbins) or past dst is dropped where the C loop
 * would write out of bounds. */
static inline int64_t hydg_weighted_histogram_i64(hyd_array_t* dst, hyd_array_t* data, int64_t nbins, int64_t n) {
    if (n <= 0 || nbins <= 0) return 0;
    hydg_check_len("hydg_weighted_histogram_i64", data, n);
    int64_t* hist = NULL;
    int64_t* host = (int64_t*)malloc((size_t)nbins * sizeof(int64_t));
    if (!host) {
        fprintf(stderr, "hydg_weighted_histogram_i64: out of memory\n");
        exit(1);
    }
    if (hydhip_malloc((void**)&hist, nbins * (int64_t)sizeof(int64_t)) != 0 ||
        hydhip_weighted_histogram_i64((const int64_t*)hyd_array_device_data(data), n, nbins, hist, 0, NULL) != 0 ||
        hydhip_memcpy_d2h(host, hist, nbins * (int64_t)sizeof(int64_t), NULL) != 0 ||
        hydhip_stream_sync(NULL) != 0 || hydhip_free(hist) != 0)
        hydg_die("hydg_weighted_histogram_i64");
    if (hyd_array_sync_host(dst) != 0) hydg_die("hydg_weighted_histogram_i64");  /* an HBM copy may be newer */
    const int64_t ndst = hyd_shape_size(dst->shape);
    uint64_t* d = (uint64_t*)dst->data;
    for (int64_t b = 0; b < nbins && b < ndst; ++b) d[b] += (uint64_t)host[b];
    free(host);
    hyd_array_mark_host_dirty(dst);
    return 0;
}

/* acc + sum over i in [0, n) of xs[i] * ys[i]: the reduction loop of
 * `reduce (+.) init (zipwith (*.) xs ys)` (CodegenC.hs:1306-1413), with the
 * accumulator's value on entry as init; the fixed tree of hydhip_dot_f64. */
static inline double hydg_dot_f64(hyd_array_t* xs, hyd_array_t* ys, int64_t n, double acc) {
    static double* out = NULL; /* one device word per program */
    if (n <= 0) return acc;
    hydg_check_len("hydg_dot_f64", xs, n);
    hydg_check_len("hydg_dot_f64", ys, n);
    if (!out && hydhip_malloc((void**)&out, (int64_t)sizeof(double)) != 0) hydg_die("hydg_dot_f64");
    double r = 0.0;
    if (hydhip_dot_f64((const double*)hyd_array_device_data(xs), (const double*)hyd_array_device_data(ys), n, acc,
                       out, NULL) != 0 ||
        hydhip_memcpy_d2h(&r, out, (int64_t)sizeof(double), NULL) != 0 || hydhip_stream_sync(NULL) != 0)
        hydg_die("hydg_dot_f64");
    return r;
}

#endif /* HYD_HIP_GLUE_H */
