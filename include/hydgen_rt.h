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
#include <string.h>

#include <algorithm>
#include <vector>

#include "hyd_exp.h"
#include "hyd_hip_runtime.h"

/* ---- shape tuples (host and device) --------------------------------------
 * The loops run over all HYD_MAX_DIMS slots with the rank as a guard, so
 * that a tuple stays in registers (a loop bound by the rank would index the
 * tuple dynamically, which puts it in scratch memory). */
__host__ __device__ static inline hyd_tuple_t hydk_tuple(int n, int64_t e0 = 0, int64_t e1 = 0, int64_t e2 = 0,
                                                         int64_t e3 = 0, int64_t e4 = 0, int64_t e5 = 0,
                                                         int64_t e6 = 0, int64_t e7 = 0) {
    hyd_tuple_t t;
    t.elems[0] = n > 0 ? e0 : 0, t.elems[1] = n > 1 ? e1 : 0, t.elems[2] = n > 2 ? e2 : 0;
    t.elems[3] = n > 3 ? e3 : 0, t.elems[4] = n > 4 ? e4 : 0, t.elems[5] = n > 5 ? e5 : 0;
    t.elems[6] = n > 6 ? e6 : 0, t.elems[7] = n > 7 ? e7 : 0;
    t.ndims = n;
    return t;
}

__host__ __device__ static inline int64_t hydk_shape_size(hyd_tuple_t s) {
    uint64_t size = 1;
#pragma unroll
    for (int i = 0; i < HYD_MAX_DIMS; ++i)
        if (i < s.ndims) size *= (uint64_t)s.elems[i];
    return (int64_t)size;
}

__host__ __device__ static inline hyd_tuple_t hydk_shape_init(hyd_tuple_t s) {
    hyd_tuple_t o = hydk_tuple(0);
    if (s.ndims <= 0) return o;
    o.ndims = s.ndims - 1;
#pragma unroll
    for (int i = 0; i < HYD_MAX_DIMS; ++i) o.elems[i] = i < o.ndims ? s.elems[i] : 0;
    return o;
}

__host__ __device__ static inline int64_t hydk_shape_last(hyd_tuple_t s) {
    int64_t last = 0;
#pragma unroll
    for (int i = 0; i < HYD_MAX_DIMS; ++i)
        if (i == s.ndims - 1) last = s.elems[i];
    return last;
}

__host__ __device__ static inline hyd_tuple_t hydk_flat_to_nd(int64_t flat, hyd_tuple_t s) {
    hyd_tuple_t idx = hydk_tuple(s.ndims);
    int64_t rem = flat;
#pragma unroll
    for (int i = HYD_MAX_DIMS - 1; i >= 0; --i)
        if (i < s.ndims) {
            idx.elems[i] = rem % s.elems[i];
            rem /= s.elems[i];
        }
    return idx;
}

__host__ __device__ static inline int64_t hydk_nd_to_flat(hyd_tuple_t nd, hyd_tuple_t s) {
    uint64_t flat = 0, stride = 1;
#pragma unroll
    for (int i = HYD_MAX_DIMS - 1; i >= 0; --i)
        if (i < s.ndims) {
            flat += (uint64_t)nd.elems[i] * stride;
            stride *= (uint64_t)s.elems[i];
        }
    return (int64_t)flat;
}

/* ---- division by a divisor that does not change over a loop ----------------
 * C's truncating q = x / m and r = x % m (CodegenC.hs:2473-2474) by
 * multiplication: Granlund & Montgomery, "Division by invariant integers
 * using multiplication" (PLDI 1994), fig. 4.1, for the unsigned |x| / |m|:
 * l = ceil(log2 |m|), magic = floor(2^64 (2^l - |m|) / |m|) + 1,
 * t = mulhi(magic, n), q = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0),
 * exact for every 0 <= n < 2^64; then the signs of C (the quotient negative
 * when exactly one operand is, the remainder the sign of x).  The launcher
 * computes the constants; m = 0 and m = INT64_MIN keep the plain operators. */
typedef struct {
    int64_t m;
    uint64_t am, magic;
    uint64_t mask;  /* |m| - 1 when |m| is a power of two (>= 2): mask and shift */
    uint32_t sh1, sh2, log2;
    int neg, fast;
} hydk_divc_t;

static inline hydk_divc_t hydk_make_divc(int64_t m) {
    hydk_divc_t d;
    d.m = m;
    d.fast = m != 0 && m != INT64_MIN;
    d.neg = m < 0;
    d.am = d.fast ? (uint64_t)(m < 0 ? -m : m) : 1;
    uint32_t l = 0;
    while (l < 64 && ((uint64_t)1 << l) < d.am) ++l;
    const unsigned __int128 num = ((unsigned __int128)(((unsigned __int128)1 << l) - d.am)) << 64;
    d.magic = (uint64_t)(num / d.am) + 1;
    d.sh1 = l < 1 ? l : 1;
    d.sh2 = l > 0 ? l - 1 : 0;
    d.mask = (d.am >= 2 && (d.am & (d.am - 1)) == 0) ? d.am - 1 : 0;
    d.log2 = l;
    return d;
}

__host__ __device__ static inline uint64_t hydk_udiv(uint64_t n, const hydk_divc_t& d) {
    if (d.mask) return n >> d.log2;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t t = __umul64hi(d.magic, n);
#else
    const uint64_t t = (uint64_t)(((unsigned __int128)d.magic * n) >> 64);
#endif
    return (t + ((n - t) >> d.sh1)) >> d.sh2;
}

__host__ __device__ static inline int64_t hydk_cdiv(int64_t x, const hydk_divc_t& d) {
    if (!d.fast) return x / d.m;
    const uint64_t n = x < 0 ? 0 - (uint64_t)x : (uint64_t)x;
    const uint64_t q = hydk_udiv(n, d);
    return ((x < 0) != (d.neg != 0)) ? (int64_t)(0 - q) : (int64_t)q;
}

__host__ __device__ static inline int64_t hydk_cmod(int64_t x, const hydk_divc_t& d) {
    if (!d.fast) return x % d.m;
    const uint64_t n = x < 0 ? 0 - (uint64_t)x : (uint64_t)x;
    const uint64_t r = d.mask ? (n & d.mask) : n - hydk_udiv(n, d) * d.am;
    return x < 0 ? (int64_t)(0 - r) : (int64_t)r;
}

/* ---- 16-byte accesses of streamed arrays -------------------------------- */
typedef double hydk_v2d __attribute__((ext_vector_type(2)));
typedef long long hydk_v2i __attribute__((ext_vector_type(2)));

static inline int hydk_aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

/* ---- device helpers ---------------------------------------------------- */

/* reduction combines: fp64 IEEE, int64 wrapping */
__device__ static inline double hydk_add_f64(double a, double b) { return a + b; }
__device__ static inline double hydk_mul_f64(double a, double b) { return a * b; }
__device__ static inline int64_t hydk_add_i64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
__device__ static inline int64_t hydk_mul_i64(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
struct hydk_add_f64_op { __device__ double operator()(double a, double b) const { return a + b; } };
struct hydk_mul_f64_op { __device__ double operator()(double a, double b) const { return a * b; } };
struct hydk_add_i64_op { __device__ int64_t operator()(int64_t a, int64_t b) const { return hydk_add_i64(a, b); } };
struct hydk_mul_i64_op { __device__ int64_t operator()(int64_t a, int64_t b) const { return hydk_mul_i64(a, b); } };

/* The value lane + D of the wave holds, D = +-1, +-2, ... (|D| wave shifts
 * through DPP: wave_shl:1 moves every value one lane down, so lane i reads
 * lane i + 1; wave_shr:1 one lane up).  Lanes with no source lane (the last
 * |D| for D > 0, the first |D| for D < 0) read 0 bits.  64-bit values move
 * as two 32-bit halves. */
template <int D, class T>
__device__ static inline T hydk_lane_shift(T v) {
    static_assert(sizeof(T) == 8, "64-bit values");
    uint64_t b = __builtin_bit_cast(uint64_t, v);
    int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
    constexpr int ctrl = D > 0 ? 0x130 : 0x138; /* wave_shl:1 / wave_shr:1 */
#pragma unroll
    for (int k = 0; k < (D > 0 ? D : -D); ++k) {
        lo = __builtin_amdgcn_update_dpp(0, lo, ctrl, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp(0, hi, ctrl, 0xf, 0xf, false);
    }
    return __builtin_bit_cast(T, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

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
 * zeroed again for the next launch on the stream.  One thread per (bin,
 * copy): `lanes` (a power of two >= used, <= 32) consecutive lanes hold the
 * copies of one bin and add them by a shuffle butterfly, so the fold costs
 * one memory round trip rather than `used` dependent ones. */
template <int D = 0>
__global__ __launch_bounds__(256) void hydk_spread_fold_t(unsigned long long* spread, int used, int lanes, int64_t ndst,
                                                          unsigned long long* dst) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = t / lanes;
    const int c = (int)(t % lanes);
    unsigned long long v = 0;
    if (b < ndst && c < used) {
        unsigned long long* q = spread + (size_t)c * ndst + b;
        v = *q;
        *q = 0;
    }
    for (int off = lanes >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (c == 0 && b < ndst) dst[b] += v;
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

/* ---- group launches --------------------------------------------------------
 * One array access of one member of a group: its bytes [lo, hi), whether
 * the member reads and writes it, and - when every write of the member lands
 * in a box of its iteration space - that box per dimension, [blo, bhi), and
 * (for more than one dimension) the shape of the flat index. */
typedef struct {
    int member;
    uintptr_t lo, hi;
    int read, written, boxed, nd;
    int64_t blo[HYD_MAX_DIMS], bhi[HYD_MAX_DIMS];
    hyd_tuple_t shape;
} hydk_acc_t;

static inline hydk_acc_t hydk_acc(int member, const void* p, int64_t bytes, int read, int written, int boxed, int nd,
                                  const int64_t* blo, const int64_t* bhi, const hyd_tuple_t* shape) {
    hydk_acc_t a;
    memset(&a, 0, sizeof a);
    a.member = member;
    a.lo = (uintptr_t)p;
    a.hi = a.lo + (uintptr_t)(bytes > 0 ? bytes : 0);
    a.read = read;
    a.written = written;
    a.boxed = boxed && nd >= 1 && nd <= HYD_MAX_DIMS;
    a.nd = nd;
    for (int d = 0; a.boxed && d < nd; ++d) {
        a.blo[d] = blo[d];
        a.bhi[d] = bhi[d];
    }
    if (shape) a.shape = *shape;
    /* more than one dimension: the flat index is one-to-one only inside the shape */
    if (a.boxed && nd > 1) {
        if (a.shape.ndims != nd) a.boxed = 0;
        for (int d = 0; a.boxed && d < nd; ++d)
            if (a.blo[d] < 0 || a.bhi[d] > a.shape.elems[d]) a.boxed = 0;
    }
    return a;
}

/* True when the members' kernels may run in one launch in any order: no
 * member reads bytes another writes, and two members write the same bytes
 * only through disjoint boxes of one array under one shape. */
static inline bool hydk_accesses_independent(const hydk_acc_t* a, int n) {
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const hydk_acc_t &x = a[i], &y = a[j];
            if (x.member == y.member) continue;
            if (x.hi <= y.lo || y.hi <= x.lo) continue; /* other bytes */
            if (!x.written && !y.written) continue;     /* both read */
            if (x.read || y.read) return false;         /* one reads what the other writes */
            if (!x.boxed || !y.boxed || x.lo != y.lo || x.nd != y.nd) return false;
            bool same_shape = x.nd == 1 || x.shape.ndims == y.shape.ndims;
            for (int d = 0; same_shape && x.nd > 1 && d < x.nd; ++d) same_shape = x.shape.elems[d] == y.shape.elems[d];
            if (!same_shape) return false;
            bool apart = false;
            for (int d = 0; d < x.nd; ++d)
                if (x.bhi[d] <= y.blo[d] || y.bhi[d] <= x.blo[d]) apart = true;
            if (!apart) return false;
        }
    return true;
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
    // per_cu 0: one tile per workgroup (the grid-stride loop then runs one
    // round), up to the largest grid one launch takes
    const int64_t cap = per_cu > 0 ? (int64_t)hydk_num_cus() * per_cu : (int64_t)0x7fffffff;
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
