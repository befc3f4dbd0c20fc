/* hydgen_rt.h - what the kernels and launchers hydgen generates include.
 *
 * Device side: the shape-tuple helpers with the runtime's semantics
 * (hydrangea_runtime.h:237-290 in the reference; hyd_shape_size and friends
 * here), wrapping integer combines, the wave64 reduction tree, the LDS and
 * global 64-bit adds of the scatters and the fold of their spread copies.
 * Host side: residency of hyd_array_t arguments (hyd_hip_runtime.h), the
 * stream the launchers share with the array uploads, the streaming grid
 * size, library scratch, and error exits that follow the reference runtime
 * (a message on stderr and exit(1)).
 *
 * Compile generated sources with -ffp-contract=off (each fp64 operation is
 * its own IEEE operation, as in the emitted C).
 */
#ifndef HYDGEN_RT_H
#define HYDGEN_RT_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>

#include "hyd_exp.h"
#include "hyd_hip_runtime.h"

/* ---- device helpers ---------------------------------------------------- */
__device__ static inline hyd_tuple_t hydk_tuple(int n, int64_t e0 = 0, int64_t e1 = 0, int64_t e2 = 0,
                                                int64_t e3 = 0, int64_t e4 = 0, int64_t e5 = 0,
                                                int64_t e6 = 0, int64_t e7 = 0) {
    hyd_tuple_t t;
    t.elems[0] = e0, t.elems[1] = e1, t.elems[2] = e2, t.elems[3] = e3;
    t.elems[4] = e4, t.elems[5] = e5, t.elems[6] = e6, t.elems[7] = e7;
    for (int i = n; i < HYD_MAX_DIMS; ++i) t.elems[i] = 0;
    t.ndims = n;
    return t;
}

__device__ static inline int64_t hydk_shape_size(hyd_tuple_t s) {
    uint64_t size = 1;
    for (int i = 0; i < s.ndims; ++i) size *= (uint64_t)s.elems[i];
    return (int64_t)size;
}

__device__ static inline hyd_tuple_t hydk_shape_init(hyd_tuple_t s) {
    hyd_tuple_t o = hydk_tuple(0);
    if (s.ndims <= 0) return o;
    o.ndims = s.ndims - 1;
    for (int i = 0; i < o.ndims; ++i) o.elems[i] = s.elems[i];
    return o;
}

__device__ static inline int64_t hydk_shape_last(hyd_tuple_t s) { return s.ndims <= 0 ? 0 : s.elems[s.ndims - 1]; }

__device__ static inline hyd_tuple_t hydk_flat_to_nd(int64_t flat, hyd_tuple_t s) {
    hyd_tuple_t idx = hydk_tuple(s.ndims);
    int64_t rem = flat;
    for (int i = s.ndims - 1; i >= 0; --i) {
        idx.elems[i] = rem % s.elems[i];
        rem /= s.elems[i];
    }
    return idx;
}

__device__ static inline int64_t hydk_nd_to_flat(hyd_tuple_t nd, hyd_tuple_t s) {
    uint64_t flat = 0, stride = 1;
    for (int i = s.ndims - 1; i >= 0; --i) {
        flat += (uint64_t)nd.elems[i] * stride;
        stride *= (uint64_t)s.elems[i];
    }
    return (int64_t)flat;
}

/* reduction combines: fp64 IEEE, int64 wrapping */
__device__ static inline double hydk_add_f64(double a, double b) { return a + b; }
__device__ static inline double hydk_mul_f64(double a, double b) { return a * b; }
__device__ static inline int64_t hydk_add_i64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
__device__ static inline int64_t hydk_mul_i64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
struct hydk_add_f64_op { __device__ double operator()(double a, double b) const { return a + b; } };
struct hydk_mul_f64_op { __device__ double operator()(double a, double b) const { return a * b; } };
struct hydk_add_i64_op { __device__ int64_t operator()(int64_t a, int64_t b) const { return hydk_add_i64(a, b); } };
struct hydk_mul_i64_op { __device__ int64_t operator()(int64_t a, int64_t b) const { return hydk_mul_i64(a, b); } };

/* xor butterfly over the 64 lanes of a wave: every lane ends with the same
 * value (each step combines two equal-length partial sums, and IEEE + and *
 * commute), so the result does not depend on which lane reads it. */
template <class T, class Op>
__device__ static inline T hydk_wave_reduce(T v, Op op) {
    for (int off = 32; off >= 1; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ static inline void hydk_lds_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ static inline void hydk_global_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* dst[b] += the `used` spread copies of bin b (mod 2^64); the copies are
 * zeroed again for the next launch on the stream. */
template <int D = 0>
__global__ __launch_bounds__(256) void hydk_spread_fold_t(unsigned long long* spread, int used, int64_t ndst,
                                                          unsigned long long* dst) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= ndst) return;
    unsigned long long s = 0;
    for (int c = 0; c < used; ++c) {
        unsigned long long* q = spread + (size_t)c * ndst + b;
        s += *q;
        *q = 0;
    }
    dst[b] += s;
}
#define hydk_spread_fold hydk_spread_fold_t<0>

/* ---- host helpers ---------------------------------------------------------- */
static inline void hydk_fail(const char* site, const char* what) {
    fprintf(stderr, "hydrangea: %s: %s\n", site, what);
    exit(1);
}

static inline void hydk_check(int rc, const char* site) {
    if (rc != 0) hydk_fail(site, hydhip_last_error());
}

static inline void hydk_check_launch(const char* site) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hydk_fail(site, hipGetErrorString(e));
}

/* The stream of the drop-in layer (the array uploads run on it too). */
static inline hipStream_t hydk_stream(void) { return (hipStream_t)hyd_hip_get_stream(); }

static inline void* hydk_device_data(hyd_array_t* a) { return hyd_array_device_data_keep(a); }
static inline void hydk_device_written(hyd_array_t* a) { hyd_array_mark_device_written(a); }

static inline int hydk_num_cus(void) {
    static int cus = 0;
    if (!cus) {
        int n = 0;
        int64_t hbm = 0, lds = 0;
        char arch[64];
        hydk_check(hydhip_device_info(&n, &hbm, &lds, arch, (int)sizeof arch), "hydgen");
        cus = n > 0 ? n : 1;
    }
    return cus;
}

/* Workgroups of a streaming grid: enough for every tile of `tile` elements,
 * at most `per_cu` per CU (grid-stride beyond that). */
static inline int64_t hydk_stream_grid(int64_t n, int64_t tile, int per_cu) {
    const int64_t tiles = (n + tile - 1) / tile;
    const int64_t cap = (int64_t)hydk_num_cus() * (per_cu > 0 ? per_cu : 1);
    return std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
}

static inline void* hydk_scratch(int slot, int64_t bytes, hipStream_t s) {
    void* p = hydhip_scratch(slot, bytes, s);
    if (!p) hydk_fail("hydgen scratch", hydhip_last_error());
    return p;
}

static inline void* hydk_zero_scratch(hipStream_t s) {
    void* p = hydhip_zero_scratch(s);
    if (!p) hydk_fail("hydgen zero scratch", hydhip_last_error());
    return p;
}

static inline void hydk_read_back(void* dst, const void* src, size_t bytes, hipStream_t s, const char* site) {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        hydk_fail(site, "reading the result back failed");
}

#endif /* HYDGEN_RT_H */
