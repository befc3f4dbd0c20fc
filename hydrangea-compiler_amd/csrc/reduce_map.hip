// reduce_map.hip - fused fp64 reduction (zipwith (*.) -> reduce (+.)) and the
// fused elementwise map chain.
//
// Reference shapes: reduce_generate lowering, src/Language/Hydrangea/
// Lowering.hs:762-836 and the step `acc = acc + elem` (Lowering.hs:3517-3535),
// emitted as `parallel for simd reduction(+:acc)` (CodegenC.hs:1306-1413);
// map/zipwith chains as one three-address statement per op
// (CodegenC.hs:2293-2400, operators :2469-2525, `if` as SIf :978-1001).
//
// CDNA4 mapping: grid-stride streaming with 16-byte loads, each thread keeps
// four independent partial sums (fixed assignment -> deterministic), then a
// wave64 xor-shuffle tree, an LDS tree across the workgroup's waves, one
// partial per workgroup and a single-workgroup fixed tree over the partials.
// Products and sums stay separate IEEE operations (-ffp-contract=off), the
// same per-element semantics as the C; only the summation order differs,
// which the 1e-12 relative bound covers.
#include <string.h>

#include <vector>

#include "../../include/hyd_exp.h"
#include "hyd_internal.hpp"

namespace hyd {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the workgroup in a fixed order; result valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[BLOCK / 64];
    v = wave_sum(v);
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) part[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        // pairwise over the (few) waves
        double a[BLOCK / 64];
#pragma unroll
        for (int i = 0; i < BLOCK / 64; ++i) a[i] = part[i];
#pragma unroll
        for (int w = 1; w < BLOCK / 64; w <<= 1)
#pragma unroll
            for (int i = 0; i + w < BLOCK / 64; i += 2 * w) a[i] = a[i] + a[i + w];
        s = a[0];
    }
    return s;
}

// Launch shape of the persistent streaming kernels: a few waves per CU, each
// thread with kStreamDepth 16-byte pairs (per array) in flight.  Measured on
// MI355X, the HBM rate of a read stream falls as the number of concurrent
// waves grows (dot: 7.3 TB/s at one 256-thread workgroup per CU, 5.8 TB/s at
// eight; the flat form below, 6.2-6.3 TB/s at 2-8 pairs per thread), so the
// dot's grid is sized in workgroups per CU, not by the data.  Streams that
// write (the map chain, the copy) run faster in the flat form.
constexpr int kStreamDepth = 4;
constexpr int kDotPerCU = 1;
constexpr int kMapPerCU = 1;
constexpr int kMapDepth = 6;
constexpr int kCopyPerCU = 1;

// Grid for `work` pairs: at most `per_cu` workgroups per CU, at least one.
inline int64_t stream_grid(int64_t work, int depth, int per_cu) {
    return std::max<int64_t>(1, std::min<int64_t>(ceil_div(work, (int64_t)kBlock * depth),
                                                  (int64_t)device_info().num_cus * per_cu));
}

template <bool VEC, int D>
__global__ __launch_bounds__(kBlock) void dot_partial_kernel(const double* __restrict__ xs,
                                                             const double* __restrict__ ys,
                                                             int64_t n, double* __restrict__ partial) {
    // four partial sums per thread; pair k of an iteration feeds a[2k % 4],
    // a[(2k + 1) % 4] (fixed assignment: deterministic)
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (VEC) {
        const int64_t npairs = n >> 1;
        for (; p + (D - 1) * step < npairs; p += D * step) {
            v2d x[D], y[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                x[k] = ld_stream2(xs + 2 * (p + k * step));
                y[k] = ld_stream2(ys + 2 * (p + k * step));
            }
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double m0 = x[k].x * y[k].x, m1 = x[k].y * y[k].y;
                a[(2 * k) & 3] = a[(2 * k) & 3] + m0;
                a[(2 * k + 1) & 3] = a[(2 * k + 1) & 3] + m1;
            }
        }
        for (; p < npairs; p += step) {
            const v2d x = ld_stream2(xs + 2 * p), y = ld_stream2(ys + 2 * p);
            const double m0 = x.x * y.x, m1 = x.y * y.y;
            a[0] = a[0] + m0;
            a[1] = a[1] + m1;
        }
        if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
            const double m = xs[n - 1] * ys[n - 1];
            a[2] = a[2] + m;
        }
    } else {
        for (; p < n; p += step) {
            const double m = xs[p] * ys[p];
            a[0] = a[0] + m;
        }
    }
    const double t = (a[0] + a[1]) + (a[2] + a[3]);
    const double s = block_sum<kBlock>(t);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

constexpr int kFinalBlock = 1024;

__global__ __launch_bounds__(kFinalBlock) void dot_final_kernel(const double* __restrict__ partial,
                                                                int nparts, double init,
                                                                double* __restrict__ out) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kFinalBlock) v = v + partial[i];
    double s = block_sum<kFinalBlock>(v);
    if (threadIdx.x == 0) *out = init + s;
}

// The fused chain of the map benchmark program (one statement per op, the
// order of the source expression).
__device__ __forceinline__ double map_chain_op(double x, double y, const uint64_t* exp_tab) {
    double t1 = 2.0 * x;
    double t2 = t1 + y;
    double t3 = __builtin_fabs(t2);
    double t4 = __builtin_sqrt(t3);
    double t5 = t4 * x;
    double t6 = t5 - y;
    double t7;
    if (t6 > 0.0)
        t7 = t6;
    else
        t7 = 0.0;
    double t8 = t7 * 1.0e-3;
    return hyd_exp_with(t8, exp_tab);
}

template <int D>
__device__ __forceinline__ void load_pairs(const double* __restrict__ xs, const double* __restrict__ ys,
                                           int64_t p, int64_t step, v2d* x, v2d* y) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
        x[k] = ld_stream2(xs + 2 * (p + k * step));
        y[k] = ld_stream2(ys + 2 * (p + k * step));
    }
}

// One workgroup per CU with 6 pairs per array in flight (measured best on
// MI355X: 6.4 TB/s; software-pipelining over two register sets, or two
// workgroups per CU, measured slower).  The exp table is copied into LDS: with the lookups as
// global loads they shared the L1 path with the stream, and the kernel then
// needed 8 workgroups per CU to reach 5.4 TB/s.
template <bool VEC, int D>
__global__ __launch_bounds__(kBlock) void map_chain_kernel(const double* __restrict__ xs,
                                                           const double* __restrict__ ys,
                                                           int64_t n, double* __restrict__ out) {
    // the exp table in LDS: its per-element lookups then stay off the path
    // of the streaming loads
    __shared__ uint64_t exp_tab[256];
    for (int i = threadIdx.x; i < 256; i += kBlock) exp_tab[i] = hyd_exp_tab_dev[i];
    __syncthreads();
    const int64_t step = (int64_t)gridDim.x * kBlock;
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (VEC) {
        const int64_t npairs = n >> 1;
        // D pairs per iteration: 2D 16-byte loads in flight before the math
        for (; p + (D - 1) * step < npairs; p += D * step) {
            v2d x[D], y[D];
            load_pairs<D>(xs, ys, p, step, x, y);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                v2d r;
                r.x = map_chain_op(x[k].x, y[k].x, exp_tab);
                r.y = map_chain_op(x[k].y, y[k].y, exp_tab);
                st_stream2(out + 2 * (p + k * step), r);
            }
        }
        for (; p < npairs; p += step) {
            const v2d x = ld_stream2(xs + 2 * p), y = ld_stream2(ys + 2 * p);
            v2d r;
            r.x = map_chain_op(x.x, y.x, exp_tab);
            r.y = map_chain_op(x.y, y.y, exp_tab);
            st_stream2(out + 2 * p, r);
        }
        if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) out[n - 1] = map_chain_op(xs[n - 1], ys[n - 1], exp_tab);
    } else {
        for (; p < n; p += step) out[p] = map_chain_op(xs[p], ys[p], exp_tab);
    }
}

// The flat form of the map chain and the copy: no loop; workgroup b takes
// the contiguous block of kBlock * K pairs at b * kBlock * K, thread t the
// pairs at t + k * kBlock in it, and the grid is sized by the data (2^27
// fp64 elements: 262144 workgroups).  Measured on MI355X against the
// persistent grid-stride form (one workgroup per CU, each thread several
// pairs a grid apart): an x + y stream 511-521 against 539-555 us, a copy
// 317-322 against 348-355 us (6.7 TB/s) - with the hardware keeping every CU
// full of short-lived waves and each workgroup's accesses one contiguous
// block, the DRAM sees more requests in flight and fewer pages open at once.
#ifndef HYD_FLAT_PAIRS
#define HYD_FLAT_PAIRS 1
#endif
constexpr int kFlatPairs = HYD_FLAT_PAIRS;

template <int K>
__global__ __launch_bounds__(kBlock) void map_chain_flat_kernel(const double* __restrict__ xs,
                                                                const double* __restrict__ ys, int64_t npairs,
                                                                int64_t n, double* __restrict__ out) {
    __shared__ uint64_t exp_tab[256];
    static_assert(kBlock == 256, "one table entry per thread");
    exp_tab[threadIdx.x] = hyd_exp_tab_dev[threadIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kBlock * K + threadIdx.x;
    v2d x[K], y[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t q = base + (int64_t)k * kBlock;
        if (q < npairs) {
            x[k] = ld_stream2(xs + 2 * q);
            y[k] = ld_stream2(ys + 2 * q);
        }
    }
    __syncthreads();  // the table is in LDS (the loads above are in flight meanwhile)
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t q = base + (int64_t)k * kBlock;
        if (q < npairs) {
            v2d r;
            r.x = map_chain_op(x[k].x, y[k].x, exp_tab);
            r.y = map_chain_op(x[k].y, y[k].y, exp_tab);
            st_stream2(out + 2 * q, r);
        }
    }
    if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        out[n - 1] = map_chain_op(xs[n - 1], ys[n - 1], exp_tab);
}

template <int K>
__global__ __launch_bounds__(kBlock) void copy_flat_kernel(const double* __restrict__ src, int64_t npairs,
                                                           double* __restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * kBlock * K + threadIdx.x;
    v2d a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t q = base + (int64_t)k * kBlock;
        if (q < npairs) a[k] = ld_stream2(src + 2 * q);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t q = base + (int64_t)k * kBlock;
        if (q < npairs) st_stream2(dst + 2 * q, a[k]);
    }
}

// Workgroups of the flat form for `npairs` pairs, or 0 when the grid would be
// too large for one launch (the grid-stride form then).
inline int64_t flat_grid(int64_t npairs) {
    const int64_t g = std::max<int64_t>(1, ceil_div(npairs, (int64_t)kBlock * kFlatPairs));
    return g <= 0x7fffffff ? g : 0;
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Plain streaming copy: the bandwidth reference the fused kernels are read
// against (and the calibration kernel for the HBM byte counters).
template <int D>
__global__ __launch_bounds__(kBlock) void copy_kernel(const double* __restrict__ src, int64_t npairs,
                                                      double* __restrict__ dst) {
    const int64_t step = (int64_t)gridDim.x * kBlock;
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; p + (D - 1) * step < npairs; p += D * step) {
        v2d a[D];
#pragma unroll
        for (int k = 0; k < D; ++k) a[k] = ld_stream2(src + 2 * (p + k * step));
#pragma unroll
        for (int k = 0; k < D; ++k) st_stream2(dst + 2 * (p + k * step), a[k]);
    }
    for (; p < npairs; p += step) st_stream2(dst + 2 * p, ld_stream2(src + 2 * p));
}

// Elementwise exp / sqrt of the map chain's math, for the parity tests of
// the math functions on their own.
__global__ void exp_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = hyd_exp(x[i]);
}

__global__ void sqrt_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = __builtin_sqrt(x[i]);
}

}  // namespace
}  // namespace hyd

using namespace hyd;

// out[i] = exp(x[i]) (op 0) or sqrt(x[i]) (op 1) on the GPU.
extern "C" int hydhip_math_f64(int op, const double* x, int64_t n, double* out, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || (n > 0 && (!x || !out)) || op < 0 || op > 1)
        return set_error(HYD_E_INVALID, "hydhip_math_f64: bad arguments");
    if (n == 0) return HYD_OK;
    int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192));
    if (op == 0)
        exp_kernel<<<(unsigned)blocks, 256, 0, resolve_stream(stream)>>>(x, n, out);
    else
        sqrt_kernel<<<(unsigned)blocks, 256, 0, resolve_stream(stream)>>>(x, n, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

// ---- batched dot ------------------------------------------------------------
// Request r is cut into fixed tiles of kDotTile elements; tile (r, t) is summed
// by one workgroup in a fixed order and request r's result is init plus its
// tile partials added in tile order.  The bits of a request therefore do not
// depend on the other requests of the batch.
namespace hyd {
namespace {
constexpr int64_t kDotTile = 65536;

struct DotTile {
    const double* xs;
    const double* ys;
    int64_t len;
    int64_t pad;
};

__global__ __launch_bounds__(kBlock) void batched_dot_tiles(const DotTile* __restrict__ tiles,
                                                           double* __restrict__ partial) {
    const DotTile t = tiles[blockIdx.x];
    double a0 = 0.0, a1 = 0.0;
    const bool vec = ((reinterpret_cast<uintptr_t>(t.xs) | reinterpret_cast<uintptr_t>(t.ys)) & 15) == 0;
    if (vec) {
        const int64_t npairs = t.len >> 1;
        for (int64_t p = threadIdx.x; p < npairs; p += kBlock) {
            v2d x = ld_stream2(t.xs + 2 * p), y = ld_stream2(t.ys + 2 * p);
            double m0 = x.x * y.x, m1 = x.y * y.y;
            a0 = a0 + m0;
            a1 = a1 + m1;
        }
        if ((t.len & 1) && threadIdx.x == 0) {
            double m = t.xs[t.len - 1] * t.ys[t.len - 1];
            a0 = a0 + m;
        }
    } else {
        for (int64_t i = threadIdx.x; i < t.len; i += kBlock) {
            double m = t.xs[i] * t.ys[i];
            a0 = a0 + m;
        }
    }
    double s = block_sum<kBlock>(a0 + a1);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void batched_dot_final(const double* __restrict__ partial, const int64_t* __restrict__ first,
                                  int64_t count, double init, double* const* __restrict__ outs) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    double acc = init;
    for (int64_t t = first[r]; t < first[r + 1]; ++t) acc = acc + partial[t];
    *outs[r] = acc;
}
}  // namespace

// Host-side plan + launches; xs/ys/outs/lens are host arrays of device
// pointers and sizes.
int batched_dot(const double* const* xs, const double* const* ys, const int64_t* lens, int64_t count,
                double init, double* const* outs, hipStream_t s) {
    std::vector<DotTile> tiles;
    std::vector<int64_t> first((size_t)count + 1);
    for (int64_t r = 0; r < count; ++r) {
        if (lens[r] < 0 || !outs[r] || (lens[r] > 0 && (!xs[r] || !ys[r])))
            return set_error(HYD_E_INVALID, "batched_dot: bad request %lld", (long long)r);
        first[r] = (int64_t)tiles.size();
        for (int64_t st = 0; st < lens[r]; st += kDotTile)
            tiles.push_back(DotTile{xs[r] + st, ys[r] + st, std::min(kDotTile, lens[r] - st), 0});
    }
    first[count] = (int64_t)tiles.size();
    const size_t tb = tiles.size() * sizeof(DotTile), fb = first.size() * 8, ob = (size_t)count * 8;
    const size_t pb = tiles.size() * 8;
    unsigned char* ws = static_cast<unsigned char*>(stream_workspace(s, tb + fb + ob + pb + 64));
    if (!ws) return set_error(HYD_E_OOM, "batched_dot: workspace allocation failed");
    HYD_TRY(stage_h2d(s, ws, tb + fb + ob, [&](unsigned char* q) {
        memcpy(q, tiles.data(), tb);
        memcpy(q + tb, first.data(), fb);
        memcpy(q + tb + fb, outs, ob);
    }));
    auto* d_tiles = reinterpret_cast<DotTile*>(ws);
    auto* d_first = reinterpret_cast<int64_t*>(ws + tb);
    auto* d_outs = reinterpret_cast<double* const*>(ws + tb + fb);
    auto* d_part = reinterpret_cast<double*>(ws + ((tb + fb + ob + 15) & ~(size_t)15));
    if (!tiles.empty()) {
        batched_dot_tiles<<<(unsigned)tiles.size(), kBlock, 0, s>>>(d_tiles, d_part);
        HYD_CHECK_LAUNCH();
    }
    batched_dot_final<<<(unsigned)ceil_div(count, 256), 256, 0, s>>>(d_part, d_first, count, init, d_outs);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

}  // namespace hyd

extern "C" int hydhip_batched_dot_f64(const double* const* xs, const double* const* ys,
                                      const int64_t* lens, int64_t count, double init,
                                      double* const* outs, void* stream) {
    HYD_TRY(ensure_init());
    if (count < 0 || (count > 0 && (!xs || !ys || !lens || !outs)))
        return set_error(HYD_E_INVALID, "hydhip_batched_dot_f64: bad arguments");
    if (count == 0) return HYD_OK;
    return batched_dot(xs, ys, lens, count, init, outs, resolve_stream(stream));
}

// dst[i] = src[i], i < n (n even, 16-byte aligned pointers).
extern "C" int hydhip_copy_f64(const double* src, int64_t n, double* dst, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || (n & 1) || (n > 0 && (!src || !dst)) || !aligned16(src) || !aligned16(dst))
        return set_error(HYD_E_INVALID, "hydhip_copy_f64: needs an even count and 16-byte aligned pointers");
    if (n == 0) return HYD_OK;
    const int64_t npairs = n >> 1;
    // untuned classes take the flat form; a tuned class its workgroups per CU
    const int per_cu = tuned_per_cu(HYD_SHAPE_COPY, n, HYD_SHAPE_FLAT);
    const int64_t flat = per_cu == HYD_SHAPE_FLAT ? flat_grid(npairs) : 0;
    if (flat) {
        copy_flat_kernel<kFlatPairs><<<(unsigned)flat, kBlock, 0, resolve_stream(stream)>>>(src, npairs, dst);
    } else {
        const int64_t blocks = stream_grid(npairs, kStreamDepth, per_cu != HYD_SHAPE_FLAT ? per_cu : kCopyPerCU);
        copy_kernel<kStreamDepth><<<(unsigned)blocks, kBlock, 0, resolve_stream(stream)>>>(src, npairs, dst);
    }
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

// The same exp compiled for the host CPU (same IEEE operation sequence).
extern "C" void hydhip_host_exp(const double* x, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = hyd_exp(x[i]);
}

using namespace hyd;

extern "C" int hydhip_dot_f64(const double* xs, const double* ys, int64_t n, double init,
                              double* out, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || !out || (n > 0 && (!xs || !ys)))
        return set_error(HYD_E_INVALID, "hydhip_dot_f64: bad arguments");
    hipStream_t s = resolve_stream(stream);
    // Fixed partial count for a given n and device: deterministic results.
    const bool vec = aligned16(xs) && aligned16(ys);
    const int64_t blocks = stream_grid(vec ? (n >> 1) : n, vec ? kStreamDepth : 1, kDotPerCU);
    double* partial = static_cast<double*>(stream_workspace(s, (size_t)blocks * sizeof(double)));
    if (!partial) return set_error(HYD_E_OOM, "hydhip_dot_f64: workspace allocation failed");
    if (vec)
        dot_partial_kernel<true, kStreamDepth><<<(unsigned)blocks, kBlock, 0, s>>>(xs, ys, n, partial);
    else
        dot_partial_kernel<false, kStreamDepth><<<(unsigned)blocks, kBlock, 0, s>>>(xs, ys, n, partial);
    HYD_CHECK_LAUNCH();
    dot_final_kernel<<<1, kFinalBlock, 0, s>>>(partial, (int)blocks, init, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

extern "C" int hydhip_map_chain_f64(const double* xs, const double* ys, int64_t n, double* out,
                                    void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || (n > 0 && (!xs || !ys || !out)))
        return set_error(HYD_E_INVALID, "hydhip_map_chain_f64: bad arguments");
    if (n == 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    const bool vec = aligned16(xs) && aligned16(ys) && aligned16(out);
    // untuned classes take the flat form; a tuned class its workgroups per CU
    const int per_cu = tuned_per_cu(HYD_SHAPE_MAP_CHAIN, n, HYD_SHAPE_FLAT);
    const int64_t flat = vec && per_cu == HYD_SHAPE_FLAT ? flat_grid(n >> 1) : 0;
    if (flat) {
        map_chain_flat_kernel<kFlatPairs><<<(unsigned)flat, kBlock, 0, s>>>(xs, ys, n >> 1, n, out);
        HYD_CHECK_LAUNCH();
        return HYD_OK;
    }
    const int64_t blocks =
        stream_grid(vec ? (n >> 1) : n, vec ? kMapDepth : 1, per_cu != HYD_SHAPE_FLAT ? per_cu : kMapPerCU);
    if (vec)
        map_chain_kernel<true, kMapDepth><<<(unsigned)blocks, kBlock, 0, s>>>(xs, ys, n, out);
    else
        map_chain_kernel<false, kMapDepth><<<(unsigned)blocks, kBlock, 0, s>>>(xs, ys, n, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}
