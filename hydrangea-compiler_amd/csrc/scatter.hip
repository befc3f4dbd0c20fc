// scatter.hip - scatter (+) into int64 bins: the weighted histogram and the
// unfused scatter-add.
//
// Reference shapes: the privatized scatter of
// src/Language/Hydrangea/CodegenC.hs:1119-1288 (per-thread private grids,
// then a per-bin merge) and the atomic scatter of CodegenC.hs:1084-1118,
// chosen by Parallelize.hs:466-503 / :350-419.  The combine is
// `new = val + old` (Lowering.hs:3553-3557); integer + is associative and
// commutative mod 2^64, so any merge order gives the emitted C's bits.
//
// CDNA4 mapping:
//  * LDS:       each workgroup keeps `copies` private copies of the bins in LDS
//               (copy = lane % copies, odd stride so copies of one bin sit in
//               different banks), adds with 64-bit LDS atomics, then merges
//               its copies into the global bins with one atomic per bin.
//  * GLOBAL:    one 64-bit global atomic per element (bins too large for LDS,
//               or too few elements to pay for a private copy).
//  * LDS_SPLIT: bins split into LDS-sized ranges; grid.y picks the range and
//               every workgroup streams all elements, keeping only its range.
#include <stdio.h>

#include <algorithm>
#include <mutex>

#include "hyd_internal.hpp"

namespace hyd {
namespace {

// route = x % bins (C truncating remainder, CodegenC.hs:2473-2474),
// value = x * x computed mod 2^64 (the C oracle is built with -fwrapv).
// POW2: bins is a power of two >= 2 (mask form of the remainder).
template <bool POW2>
struct ModSquare {
    ModConst mod;
    __device__ __forceinline__ void operator()(int64_t x, int64_t& b, uint64_t& v) const {
        if (POW2) {
            // m = 2^k <= 2^31: the low bits r in 32-bit arithmetic; a negative x
            // with r != 0 has the negative C remainder r - m, which is no bin
            const uint32_t r = (uint32_t)x & (uint32_t)mod.mask;
            b = (x >= 0 || r == 0) ? (int64_t)r : -1;
        } else {
            b = c_mod(x, mod);
        }
        v = (uint64_t)x * (uint64_t)x;
    }
};

// 16-byte load of a pair: non-temporal when the stream is read once (plain
// loads of a read-once stream keep the CU's L1 allocating and stall its
// address path - measured on MI355X), plain when several workgroups read
// the same elements and should find them in L2.
template <bool ONCE>
__device__ __forceinline__ v2i load_pair(const int64_t* p) {
    if (ONCE) return ld_stream2(p);
    return *reinterpret_cast<const v2i*>(p);
}

// Fused source: one int64 stream produces both route and value.
template <bool POW2, bool ONCE = true>
struct FusedSrc {
    const int64_t* __restrict__ data;
    ModSquare<POW2> fn;
    FusedSrc<POW2, false> shared() const { return FusedSrc<POW2, false>{data, fn}; }
    __device__ __forceinline__ void pair(int64_t p, int64_t& b0, uint64_t& v0, int64_t& b1,
                                         uint64_t& v1) const {
        const v2i x = load_pair<ONCE>(data + 2 * p);
        fn(x.x, b0, v0);
        fn(x.y, b1, v1);
    }
    __device__ __forceinline__ void one(int64_t i, int64_t& b, uint64_t& v) const {
        fn(data[i], b, v);
    }
};

// Unfused source: separate route and value arrays.
template <bool ONCE = true>
struct PairSrc {
    const int64_t* __restrict__ idx;
    const int64_t* __restrict__ vals;
    PairSrc<false> shared() const { return PairSrc<false>{idx, vals}; }
    __device__ __forceinline__ void pair(int64_t p, int64_t& b0, uint64_t& v0, int64_t& b1,
                                         uint64_t& v1) const {
        const v2i i2 = load_pair<ONCE>(idx + 2 * p), v2 = load_pair<ONCE>(vals + 2 * p);
        b0 = i2.x;
        b1 = i2.y;
        v0 = (uint64_t)v2.x;
        v1 = (uint64_t)v2.y;
    }
    __device__ __forceinline__ void one(int64_t i, int64_t& b, uint64_t& v) const {
        b = idx[i];
        v = (uint64_t)vals[i];
    }
};

// Guarded unfused source (scatter_guarded, Lowering.hs:1432-1517): element i
// contributes vals[i] at idx[i] when guard[i] != 0 (bools are int64 in the
// emitted C); a false guard routes to -1, which every form drops.
template <bool ONCE = true>
struct GuardedPairSrc {
    const int64_t* __restrict__ idx;
    const int64_t* __restrict__ vals;
    const int64_t* __restrict__ guard;
    GuardedPairSrc<false> shared() const { return GuardedPairSrc<false>{idx, vals, guard}; }
    __device__ __forceinline__ void pair(int64_t p, int64_t& b0, uint64_t& v0, int64_t& b1,
                                         uint64_t& v1) const {
        const v2i i2 = load_pair<ONCE>(idx + 2 * p), v2 = load_pair<ONCE>(vals + 2 * p),
                  g2 = load_pair<ONCE>(guard + 2 * p);
        b0 = g2.x ? i2.x : -1;
        b1 = g2.y ? i2.y : -1;
        v0 = (uint64_t)v2.x;
        v1 = (uint64_t)v2.y;
    }
    __device__ __forceinline__ void one(int64_t i, int64_t& b, uint64_t& v) const {
        b = guard[i] ? idx[i] : -1;
        v = (uint64_t)vals[i];
    }
};

__device__ __forceinline__ void lds_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void global_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Streams the elements assigned to this workgroup (grid-stride over 16-byte
// pairs, kDepth pairs in flight per thread: with one 512-thread workgroup per
// CU that is 32 KiB of loads in flight per CU - measured on MI355X as good as
// deeper pipelines) and calls add(bin, value).
constexpr int kDepth = 4;

template <int BLOCK, bool VEC, class Src, class Add>
__device__ __forceinline__ void stream_elements(const Src& src, int64_t n, Add add) {
    const int64_t step = (int64_t)gridDim.x * BLOCK;
    int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (VEC) {
        const int64_t npairs = n >> 1;
        for (; p + (kDepth - 1) * step < npairs; p += kDepth * step) {
            int64_t b[2 * kDepth];
            uint64_t v[2 * kDepth];
#pragma unroll
            for (int k = 0; k < kDepth; ++k)
                src.pair(p + k * step, b[2 * k], v[2 * k], b[2 * k + 1], v[2 * k + 1]);
#pragma unroll
            for (int k = 0; k < 2 * kDepth; ++k) add(b[k], v[k]);
        }
        for (; p < npairs; p += step) {
            int64_t b0, b1;
            uint64_t v0, v1;
            src.pair(p, b0, v0, b1, v1);
            add(b0, v0);
            add(b1, v1);
        }
        if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
            int64_t b;
            uint64_t v;
            src.one(n - 1, b, v);
            add(b, v);
        }
    } else {
        for (; p < n; p += step) {
            int64_t b;
            uint64_t v;
            src.one(p, b, v);
            add(b, v);
        }
    }
}

// LDS-private bins.  Merge: every workgroup adds its bins into spread copy
// (blockIdx.x % ncopy) of a zeroed scratch block, so each scratch address
// sees gridDim/ncopy atomics instead of gridDim; spread_fold_kernel then sums
// the copies into dst and zeroes them again (a second launch: a per-workgroup
// device-scope fence to find the last workgroup costs far more than it).
template <int BLOCK, bool VEC, class Src>
__global__ __launch_bounds__(BLOCK) void scatter_lds_kernel(Src src, int64_t n, int nbins,
                                                            int copies, int stride, int ncopy,
                                                            unsigned long long* __restrict__ spread) {
    extern __shared__ unsigned long long lds[];
    const int total = copies * stride;
    for (int i = threadIdx.x; i < total; i += BLOCK) lds[i] = 0;
    __syncthreads();
    unsigned long long* mine = lds + (threadIdx.x % copies) * stride;
    const uint64_t ub = (uint64_t)nbins;
    stream_elements<BLOCK, VEC>(src, n, [&](int64_t b, uint64_t v) {
        if ((uint64_t)b < ub) lds_add(&mine[b], v);
    });
    __syncthreads();
    unsigned long long* my_spread = spread + (size_t)(blockIdx.x % ncopy) * nbins;
    for (int b = threadIdx.x; b < nbins; b += BLOCK) {
        unsigned long long s = 0;
        for (int c = 0; c < copies; ++c) s += lds[c * stride + b];
        if (s) global_add(&my_spread[b], s);
    }
}

// dst[b] = (accumulate ? dst[b] : 0) + sum of the `used` spread copies of bin
// b; the copies are zeroed for the next launch.  One thread per (bin, copy):
// `lanes` (a power of two >= used, <= 64) consecutive lanes hold the copies of
// one bin and add them with a shuffle butterfly, so the whole fold costs one
// memory round trip instead of `used` dependent ones.
constexpr int kFoldBlock = 256;

__global__ __launch_bounds__(kFoldBlock) void spread_fold_kernel(unsigned long long* __restrict__ spread,
                                                                 int used, int lanes, int nbins,
                                                                 int accumulate,
                                                                 unsigned long long* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * kFoldBlock + threadIdx.x;
    const int64_t b = t / lanes;
    const int c = (int)(t % lanes);
    unsigned long long v = 0;
    if (b < nbins && c < used) {
        unsigned long long* p = spread + (size_t)c * nbins + b;
        v = *p;
        *p = 0;
    }
    for (int off = lanes >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (c == 0 && b < nbins) dst[b] = accumulate ? dst[b] + v : v;
}

template <int BLOCK, bool VEC, class Src>
__global__ __launch_bounds__(BLOCK) void scatter_split_kernel(Src src, int64_t n, int64_t nbins,
                                                              int range,
                                                              unsigned long long* __restrict__ dst) {
    extern __shared__ unsigned long long lds[];
    const int64_t lo = (int64_t)blockIdx.y * range;
    const int64_t hi = min(lo + (int64_t)range, nbins);
    const int width = (int)(hi - lo);
    for (int i = threadIdx.x; i < width; i += BLOCK) lds[i] = 0;
    __syncthreads();
    stream_elements<BLOCK, VEC>(src, n, [&](int64_t b, uint64_t v) {
        if ((uint64_t)(b - lo) < (uint64_t)width) lds_add(&lds[b - lo], v);
    });
    __syncthreads();
    for (int b = threadIdx.x; b < width; b += BLOCK) {
        unsigned long long s = lds[b];
        if (s) global_add(&dst[lo + b], s);
    }
}

template <int BLOCK, bool VEC, class Src>
__global__ __launch_bounds__(BLOCK) void scatter_global_kernel(Src src, int64_t n, int64_t nbins,
                                                               unsigned long long* __restrict__ dst) {
    const uint64_t ub = (uint64_t)nbins;
    stream_elements<BLOCK, VEC>(src, n, [&](int64_t b, uint64_t v) {
        if ((uint64_t)b < ub) global_add(&dst[b], v);
    });
}

constexpr int kBlock = 512;

// Lets kernel K take up to `bytes` of dynamic LDS: set once per kernel (the
// first call's bytes must be the largest any launch asks for), not per
// launch - the attribute call costs host time on every histogram otherwise.
template <auto K>
hipError_t allow_lds(size_t bytes) {
    static const hipError_t e =
        hipFuncSetAttribute((const void*)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return e;
}
constexpr size_t kLdsTarget = 64 * 1024;  // per workgroup: leaves room for 2 per CU

// The LDS form's merge scratch: by default the stream's zeroed scratch
// block; `words` long at `ptr` when a caller supplies its own (cleared by
// the call before use when `clear` is set).
struct MergeScratch {
    unsigned long long* ptr = nullptr;
    int64_t words = 0;
    bool clear = false;
};

template <class Src>
int launch_scatter(const Src& src, bool vec, int64_t n, int64_t nbins, int64_t* dst_i64,
                   int strategy, bool zero_first, hipStream_t s, MergeScratch own = MergeScratch()) {
    const DeviceInfo& di = device_info();
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(dst_i64);
    // at most 128 KiB of the 160 KiB LDS per workgroup (power-of-two bin ranges)
    const size_t lds_cap = std::min(di.lds_per_block, (size_t)128 * 1024);
    const int64_t per_block_elems = (int64_t)kBlock * 8;
    int64_t want_blocks = std::max<int64_t>(1, ceil_div(n, per_block_elems));

    if (strategy == HYD_SCATTER_AUTO) {
        // Size rule: LDS bins when a private copy fits and there are >= 16
        // elements per bin to pay for zeroing and merging it, the split form
        // for large bin counts with many elements, else global atomics.
        int64_t stride = nbins | 1;
        bool fits = (size_t)stride * 8 <= lds_cap;
        // a supplied merge scratch must hold the one spread copy of larger bins
        if (own.ptr && nbins > own.words) fits = false;
        if (fits && n >= 16 * nbins)
            strategy = HYD_SCATTER_LDS;
        else if (!fits && n >= 8 * nbins)
            strategy = HYD_SCATTER_LDS_SPLIT;
        else
            strategy = HYD_SCATTER_GLOBAL;
    }

    if (strategy == HYD_SCATTER_LDS) {
        int stride = (int)(nbins | 1);
        if ((size_t)stride * 8 > lds_cap)
            return set_error(HYD_E_INVALID, "scatter: %lld bins do not fit LDS (%zu bytes)",
                             (long long)nbins, lds_cap);
        int copies = (int)std::max<size_t>(1, std::min<size_t>(16, kLdsTarget / ((size_t)stride * 8)));
        size_t lds = (size_t)copies * stride * 8;
        // One workgroup per CU streams at full HBM rate (measured: the main
        // loop takes the same time from 256 to 2048 workgroups) and keeps
        // the merge atomics few.
        int64_t blocks = std::min<int64_t>(want_blocks, (int64_t)di.num_cus);
        // spread copies in the zeroed scratch: 4096 words' worth, <= 32
        int ncopy = (int)std::max<int64_t>(1, std::min<int64_t>(32, 4096 / nbins));
        unsigned long long* spread = nullptr;
        if (own.ptr) {
            ncopy = (int)std::max<int64_t>(1, std::min<int64_t>(ncopy, own.words / nbins));
            if ((int64_t)ncopy * nbins > own.words)
                return set_error(HYD_E_INVALID,
                                 "scatter: %lld bins exceed the supplied merge scratch of %lld words",
                                 (long long)nbins, (long long)own.words);
            spread = own.ptr;
            if (own.clear) HYD_TRY(fill_zero(spread, (size_t)ncopy * nbins * 8, s));
        } else {
            unsigned char* zs = static_cast<unsigned char*>(stream_zero_scratch(s));
            if (!zs) return set_error(HYD_E_OOM, "scatter: scratch allocation failed");
            if ((size_t)ncopy * nbins * 8 > kZeroScratchBytes)
                return set_error(HYD_E_INVALID, "scatter: %lld bins exceed the merge scratch",
                                 (long long)nbins);
            spread = reinterpret_cast<unsigned long long*>(zs);
        }
        int accumulate = zero_first ? 0 : 1;
        if (vec) {
            HYD_TRY_HIP((allow_lds<scatter_lds_kernel<kBlock, true, Src>>(lds_cap)));
            scatter_lds_kernel<kBlock, true, Src><<<(unsigned)blocks, kBlock, lds, s>>>(src, n, (int)nbins, copies,
                                                                                        stride, ncopy, spread);
        } else {
            HYD_TRY_HIP((allow_lds<scatter_lds_kernel<kBlock, false, Src>>(lds_cap)));
            scatter_lds_kernel<kBlock, false, Src><<<(unsigned)blocks, kBlock, lds, s>>>(src, n, (int)nbins, copies,
                                                                                         stride, ncopy, spread);
        }
        HYD_CHECK_LAUNCH();
        const int used = (int)std::min<int64_t>(ncopy, blocks);
        int lanes = 1;
        while (lanes < used) lanes <<= 1;  // <= 32
        spread_fold_kernel<<<(unsigned)ceil_div(nbins * lanes, kFoldBlock), kFoldBlock, 0, s>>>(
            spread, used, lanes, (int)nbins, accumulate, dst);
        HYD_CHECK_LAUNCH();
        return HYD_OK;
    }
    // the other forms accumulate into dst
    if (zero_first) HYD_TRY(fill_zero(dst, (size_t)nbins * 8, s));
    if (strategy == HYD_SCATTER_LDS) {
        // unreachable (handled above)
    } else if (strategy == HYD_SCATTER_LDS_SPLIT) {
        int range = (int)std::min<int64_t>(nbins, (int64_t)(lds_cap / 8));
        int64_t nranges = ceil_div(nbins, range);
        if (nranges > 65535) return set_error(HYD_E_INVALID, "scatter: too many bin ranges");
        size_t lds = (size_t)range * 8;
        int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / lds));
        int64_t bx = std::max<int64_t>(1, std::min<int64_t>(want_blocks,
                                                            (int64_t)di.num_cus * per_cu / nranges));
        dim3 grid((unsigned)bx, (unsigned)nranges);
        // every bin range streams all elements: plain loads, shared through L2
        auto ssrc = src.shared();
        using SSrc = decltype(ssrc);
        if (vec) {
            HYD_TRY_HIP((allow_lds<scatter_split_kernel<kBlock, true, SSrc>>(lds_cap)));
            scatter_split_kernel<kBlock, true, SSrc><<<grid, kBlock, lds, s>>>(ssrc, n, nbins, range, dst);
        } else {
            HYD_TRY_HIP((allow_lds<scatter_split_kernel<kBlock, false, SSrc>>(lds_cap)));
            scatter_split_kernel<kBlock, false, SSrc><<<grid, kBlock, lds, s>>>(ssrc, n, nbins, range, dst);
        }
    } else if (strategy == HYD_SCATTER_GLOBAL) {
        int64_t blocks = std::min<int64_t>(want_blocks, (int64_t)di.num_cus * 8);
        if (vec)
            scatter_global_kernel<kBlock, true, Src><<<(unsigned)blocks, kBlock, 0, s>>>(src, n, nbins, dst);
        else
            scatter_global_kernel<kBlock, false, Src><<<(unsigned)blocks, kBlock, 0, s>>>(src, n, nbins, dst);
    } else {
        return set_error(HYD_E_INVALID, "scatter: unknown strategy %d", strategy);
    }
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace
}  // namespace hyd

// ---- the no-allocation histogram's merge scratch --------------------------
// A file-scope device array, one copy per GPU the library runs on, held until
// the process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: host code uses the per-device address that
// noalloc_scratch_setup() looks up, never the name.
constexpr int64_t kNoAllocScratchWords = 8192;  // 64 KiB: bins up to 8192 in the LDS form
static __device__ __align__(256) unsigned long long scratch[kNoAllocScratchWords];

namespace hyd {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned long long* g_noalloc_ptr[kMaxDevices];  // device address of `scratch`, per device
}  // namespace

// Looks up the device address of `scratch` on `device` (which must be the
// current device) once; later calls return at once.
int noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        return set_error(HYD_E_INVALID, "no-allocation scratch: device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return HYD_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(HYD_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                         hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<unsigned long long*>(p);
    return HYD_OK;
}
}  // namespace hyd

using namespace hyd;

// hydhip_weighted_histogram_i64 with the LDS form's merge scratch in the
// file-scope `scratch` array instead of the stream's allocated block: no
// device memory is allocated by this call, nor by the library on its behalf
// (the library must have been initialised: hyd_hip_init or any earlier call).
extern "C" int hydhip_weighted_histogram_i64_noalloc(const int64_t* data, int64_t n, int64_t nbins,
                                                     int64_t* hist, int strategy, void* stream) {
    if (!library_initialized())
        return set_error(HYD_E_INVALID,
                         "hydhip_weighted_histogram_i64_noalloc: the library is not initialised "
                         "(call hyd_hip_init first: this entry point performs no set-up)");
    if (n < 0 || nbins < 1 || nbins > (int64_t)1 << 31 || !hist || (n > 0 && !data))
        return set_error(HYD_E_INVALID,
                         "hydhip_weighted_histogram_i64_noalloc: bad arguments (n=%lld nbins=%lld)",
                         (long long)n, (long long)nbins);
    if (strategy == HYD_SCATTER_LDS && nbins > kNoAllocScratchWords)
        return set_error(HYD_E_INVALID,
                         "hydhip_weighted_histogram_i64_noalloc: the LDS form's merge needs %lld words of "
                         "scratch, more than the %lld the no-allocation scratch holds",
                         (long long)nbins, (long long)kNoAllocScratchWords);
    int dev = 0;
    HYD_TRY_HIP(hipGetDevice(&dev));
    HYD_TRY(noalloc_scratch_setup(dev));  // done at initialisation for the library's device
    hipStream_t s = resolve_stream(stream);
    if (n == 0) return fill_zero(hist, (size_t)nbins * 8, s);
    const MergeScratch own{g_noalloc_ptr[dev], kNoAllocScratchWords, true};
    const ModConst mod = make_mod_const(nbins);
    if (mod.mask)
        return launch_scatter(FusedSrc<true>{data, ModSquare<true>{mod}}, aligned16(data), n, nbins,
                              hist, strategy, true, s, own);
    return launch_scatter(FusedSrc<false>{data, ModSquare<false>{mod}}, aligned16(data), n, nbins,
                          hist, strategy, true, s, own);
}

extern "C" int hydhip_weighted_histogram_i64(const int64_t* data, int64_t n, int64_t nbins,
                                             int64_t* hist, int strategy, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || nbins < 1 || nbins > (int64_t)1 << 31 || !hist || (n > 0 && !data))
        return set_error(HYD_E_INVALID,
                         "hydhip_weighted_histogram_i64: bad arguments (n=%lld nbins=%lld)",
                         (long long)n, (long long)nbins);
    hipStream_t s = resolve_stream(stream);
    if (n == 0) return fill_zero(hist, (size_t)nbins * 8, s);  // fill [nbins] 0
    // fill [nbins] 0 semantics: the LDS form writes every bin, the others
    // zero the destination first (with a kernel, so graph capture works)
    const ModConst mod = make_mod_const(nbins);
    if (mod.mask)
        return launch_scatter(FusedSrc<true>{data, ModSquare<true>{mod}}, aligned16(data), n, nbins,
                              hist, strategy, true, s);
    return launch_scatter(FusedSrc<false>{data, ModSquare<false>{mod}}, aligned16(data), n, nbins,
                          hist, strategy, true, s);
}

extern "C" int hydhip_scatter_add_i64(const int64_t* idx, const int64_t* vals, int64_t n,
                                      int64_t* dst, int64_t ndst, int strategy, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || ndst < 1 || ndst > (int64_t)1 << 31 || !dst || (n > 0 && (!idx || !vals)))
        return set_error(HYD_E_INVALID, "hydhip_scatter_add_i64: bad arguments");
    if (n == 0) return HYD_OK;
    PairSrc<> src{idx, vals};
    return launch_scatter(src, aligned16(idx) && aligned16(vals), n, ndst, dst, strategy, false,
                          resolve_stream(stream));
}

// Guarded scatter (+) into dst (which holds the defaults): dst[idx[i]] +=
// vals[i] where guard[i] != 0; guard NULL: every element.  int64 + is
// order-free mod 2^64, so the privatized forms give the serial loop's bits.
extern "C" int hydhip_scatter_guarded_add_i64(const int64_t* idx, const int64_t* vals,
                                              const int64_t* guard, int64_t n, int64_t* dst,
                                              int64_t ndst, int strategy, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || ndst < 1 || ndst > (int64_t)1 << 31 || !dst || (n > 0 && (!idx || !vals)))
        return set_error(HYD_E_INVALID, "hydhip_scatter_guarded_add_i64: bad arguments");
    if (n == 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    if (!guard)
        return launch_scatter(PairSrc<>{idx, vals}, aligned16(idx) && aligned16(vals), n, ndst, dst,
                              strategy, false, s);
    return launch_scatter(GuardedPairSrc<>{idx, vals, guard},
                          aligned16(idx) && aligned16(vals) && aligned16(guard), n, ndst, dst, strategy,
                          false, s);
}

// Scatter chain (EScatterChain, Lowering.hs:1622-1858) with (+) on int64: the
// phases accumulate into one buffer in turn.  idx/vals/guard/lens are host
// arrays of device pointers and lengths (guard[k] NULL: phase k unguarded).
extern "C" int hydhip_scatter_chain_add_i64(int64_t nphases, const int64_t* const* idx,
                                            const int64_t* const* vals, const int64_t* const* guard,
                                            const int64_t* lens, int64_t* dst, int64_t ndst,
                                            int strategy, void* stream) {
    HYD_TRY(ensure_init());
    if (nphases < 0 || (nphases > 0 && (!idx || !vals || !lens)))
        return set_error(HYD_E_INVALID, "hydhip_scatter_chain_add_i64: bad arguments");
    for (int64_t k = 0; k < nphases; ++k)
        HYD_TRY(hydhip_scatter_guarded_add_i64(idx[k], vals[k], guard ? guard[k] : nullptr, lens[k], dst,
                                               ndst, strategy, stream));
    return HYD_OK;
}

// The remainder the kernels route with, compiled for the host CPU.
extern "C" int hydhip_host_mod_i64(const int64_t* x, int64_t n, int64_t m, int64_t* out) {
    if (n < 0 || m < 1 || (n > 0 && (!x || !out)))
        return set_error(HYD_E_INVALID, "hydhip_host_mod_i64: bad arguments");
    const ModConst k = make_mod_const(m);
    for (int64_t i = 0; i < n; ++i) out[i] = k.mask ? c_mod_pow2(x[i], k) : c_mod(x[i], k);
    return HYD_OK;
}
