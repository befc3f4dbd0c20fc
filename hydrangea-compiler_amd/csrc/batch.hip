// batch.hip - many uneven weighted-histogram instances in one persistent
// launch, with the privatization strategy picked per instance size class.
//
// Each instance k is the program of programs/weighted_histogram.hyd with its
// own bin count: hist_k[x % bins_k] += x*x over data_k (Parallelize.hs:466-503
// privatizes every one of them on the CPU).  On the GPU the host cuts the
// batch into tiles, largest first, and persistent workgroups pull them from an
// atomic queue counter.  A tile is (instance, element range, bin range, mode):
//   * OWNER:   the workgroup holds the tile's bin range in LDS (several copies
//              for small bin counts), streams every element of the instance and
//              writes the bin range with plain stores - no atomics, no zeroing.
//              Instances with more bins than LDS holds get one tile per
//              16384-bin range (each streams the instance once).
//   * CHUNK:   like OWNER for one element chunk of a long instance; the chunks
//              merge their LDS bins into the (pre-zeroed) output with one
//              global atomic per touched bin.
//   * GLOBAL:  one 64-bit global atomic per element into the pre-zeroed output.
// Global atomics on gfx950 run at tens of G/s, so OWNER is preferred; the
// autotuner (tune.hip) can still pick GLOBAL per (bin class, size class).
// Integer + is order-free mod 2^64, so every schedule gives the oracle's bits.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "hyd_internal.hpp"

namespace hyd {

// Autotuner interface (tune.hip): 0 = LDS tiles, 1 = global atomics, -1 untuned.
int tuned_batch_mode(int64_t n, int64_t bins);

namespace {

enum TileMode : int32_t { kModeOwner = 0, kModeChunk = 1, kModeGlobal = 2 };

struct Tile {
    int64_t inst;
    int64_t start;  // first element (relative to the instance)
    int64_t len;
    int32_t lo;     // first bin of the range
    int32_t width;  // bins in the range
    int32_t mode;
    int32_t pad;
};

struct InstDesc {
    const int64_t* src;       // the instance's elements
    unsigned long long* dst;  // its bins
    int64_t bins;
    ModConst mod;             // route = x % bins
};

// Two launch shapes: bins <= 4096 (32 KiB LDS, 4 workgroups per CU) and bin
// ranges of 16384 (128 KiB LDS, one 1024-thread workgroup per CU).
constexpr int kSmallBlock = 512;
constexpr int kSmallWords = 4097;
constexpr int kLargeBlock = 1024;
constexpr int kLargeWords = 16384;
constexpr int64_t kChunk = (int64_t)1 << 18;  // elements per tile of a long instance

__device__ __forceinline__ void lds_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Calls add(x) for every element of p[0..len) handled by this workgroup's
// threads (16-byte loads when aligned, 4 pairs in flight per thread).
template <int BLOCK, class Add>
__device__ __forceinline__ void tile_elements(const int64_t* __restrict__ p, int64_t len, Add add) {
    if (len <= 0) return;
    if (reinterpret_cast<uintptr_t>(p) & 15) {
        if (threadIdx.x == 0) add(p[0]);
        ++p;
        --len;
    }
    const longlong2* p2 = reinterpret_cast<const longlong2*>(p);
    const int64_t npairs = len >> 1;
    int64_t q = threadIdx.x;
    for (; q + 3 * BLOCK < npairs; q += 4 * BLOCK) {
        longlong2 a = p2[q], b = p2[q + BLOCK], c = p2[q + 2 * BLOCK], d = p2[q + 3 * BLOCK];
        add(a.x); add(a.y); add(b.x); add(b.y);
        add(c.x); add(c.y); add(d.x); add(d.y);
    }
    for (; q < npairs; q += BLOCK) {
        longlong2 a = p2[q];
        add(a.x);
        add(a.y);
    }
    if ((len & 1) && threadIdx.x == 0) add(p[len - 1]);
}

template <int BLOCK, int WORDS>
__global__ __launch_bounds__(BLOCK) void batched_hist_kernel(const InstDesc* __restrict__ inst,
                                                             const Tile* __restrict__ tiles,
                                                             int64_t ntiles,
                                                             unsigned int* __restrict__ counter) {
    extern __shared__ unsigned long long lds[];
    __shared__ int64_t cur;
    for (;;) {
        if (threadIdx.x == 0) cur = (int64_t)atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t t = cur;
        __syncthreads();  // everyone has read `cur` before thread 0 may overwrite it
        if (t >= ntiles) break;  // uniform: every wave of the workgroup leaves
        const Tile tile = tiles[t];
        const InstDesc d = inst[tile.inst];
        const int64_t nb = d.bins;
        unsigned long long* h = d.dst;
        const int64_t* src = d.src + tile.start;
        // the remainder's form is chosen per tile, never per element (a branch
        // per element would serialise the loads of the streaming loop)
        const bool pow2 = d.mod.mask != 0;
        if (tile.mode == kModeGlobal) {
            auto add = [&](int64_t b, int64_t x) {
                if ((uint64_t)b < (uint64_t)nb) global_add(&h[b], (uint64_t)x * (uint64_t)x);
            };
            if (pow2)
                tile_elements<BLOCK>(src, tile.len, [&](int64_t x) { add(c_mod_pow2(x, d.mod), x); });
            else
                tile_elements<BLOCK>(src, tile.len, [&](int64_t x) { add(c_mod(x, d.mod), x); });
            continue;
        }
        const int64_t lo = tile.lo;
        const int width = tile.width;
        // copies of the range at an odd stride (a bin's copies in different
        // banks); a range that fills LDS gets one copy at stride = width
        // (the planner keeps width <= WORDS)
        int stride = width | 1;
        int copies = WORDS / stride;
        if (copies < 1) {
            copies = 1;
            stride = width;
        } else if (copies > 16) {
            copies = 16;
        }
        const int words = copies * stride;
        for (int i = threadIdx.x; i < words; i += BLOCK) lds[i] = 0;
        __syncthreads();
        unsigned long long* mine = lds + (threadIdx.x % copies) * stride;
        auto add = [&](int64_t b, int64_t x) {
            b -= lo;
            if ((uint64_t)b < (uint64_t)width) lds_add(&mine[b], (uint64_t)x * (uint64_t)x);
        };
        if (pow2)
            tile_elements<BLOCK>(src, tile.len, [&](int64_t x) { add(c_mod_pow2(x, d.mod), x); });
        else
            tile_elements<BLOCK>(src, tile.len, [&](int64_t x) { add(c_mod(x, d.mod), x); });
        __syncthreads();
        for (int b = threadIdx.x; b < width; b += BLOCK) {
            unsigned long long s = 0;
            for (int c = 0; c < copies; ++c) s += lds[c * stride + b];
            if (tile.mode == kModeOwner)
                h[lo + b] = s;
            else if (s)
                global_add(&h[lo + b], s);
        }
        __syncthreads();  // LDS reused by the next tile
    }
}

// Zeroes the bins of the instances whose tiles accumulate ((pointer, length)
// pairs).
__global__ void zero_ranges_kernel(const int64_t* __restrict__ ranges, int64_t nranges) {
    for (int64_t r = blockIdx.x; r < nranges; r += gridDim.x) {
        unsigned long long* p = reinterpret_cast<unsigned long long*>(ranges[2 * r]);
        const int64_t len = ranges[2 * r + 1];
        for (int64_t i = threadIdx.x; i < len; i += blockDim.x) p[i] = 0;
    }
}

// Size rule for classes the autotuner has not measured: LDS tiles unless the
// instance is so small next to its bins that zeroing them costs more than
// the per-element atomics.
int rule_mode(int64_t n, int64_t bins) { return n * 8 >= bins ? 0 : 1; }

std::mutex g_plan_mu;

}  // namespace

// Instance k: src[k][0..len[k]) -> dst[k][0..bins[k]).  All arrays are host
// arrays of device pointers / sizes.
int batched_histogram_ptrs(const int64_t* const* src, const int64_t* len, const int64_t* bins,
                           int64_t* const* dst, int64_t count, int policy, hipStream_t s,
                           int64_t* out_tiles) {
    std::vector<InstDesc> inst((size_t)count);
    std::vector<Tile> small, large;
    std::vector<int64_t> zero;  // (dst pointer, bins) pairs
    for (int64_t k = 0; k < count; ++k) {
        if (len[k] < 0 || bins[k] < 1 || bins[k] > ((int64_t)1 << 31) || !dst[k] || (len[k] > 0 && !src[k]))
            return set_error(HYD_E_INVALID, "batched_histogram: bad descriptor %lld", (long long)k);
        const int64_t n = len[k], nb = bins[k];
        inst[k] = InstDesc{src[k], reinterpret_cast<unsigned long long*>(dst[k]), nb, make_mod_const(nb)};
        int mode;
        if (policy == 1)
            mode = 0;
        else if (policy == 2)
            mode = 1;
        else {
            mode = tuned_batch_mode(n, nb);
            if (mode < 0) mode = rule_mode(n, nb);
        }
        if (mode == 1) {
            zero.push_back((int64_t)reinterpret_cast<uintptr_t>(dst[k]));
            zero.push_back(nb);
            for (int64_t st = 0; st < n; st += kChunk)
                small.push_back(Tile{k, st, std::min(kChunk, n - st), 0, 0, kModeGlobal, 0});
            continue;
        }
        const bool fits_small = nb <= kSmallWords - 1;
        const int64_t range = fits_small ? nb : kLargeWords;
        const bool chunked = n > kChunk;
        if (chunked) {
            zero.push_back((int64_t)reinterpret_cast<uintptr_t>(dst[k]));
            zero.push_back(nb);
        }
        std::vector<Tile>& dst = fits_small ? small : large;
        for (int64_t lo = 0; lo < nb; lo += range) {
            int32_t w = (int32_t)std::min(range, nb - lo);
            if (!chunked) {
                dst.push_back(Tile{k, 0, n, (int32_t)lo, w, kModeOwner, 0});
            } else {
                for (int64_t st = 0; st < n; st += kChunk)
                    dst.push_back(Tile{k, st, std::min(kChunk, n - st), (int32_t)lo, w, kModeChunk, 0});
            }
        }
    }
    // largest first (longest-processing-time order for the queue)
    auto by_len = [](const Tile& a, const Tile& b) { return a.len > b.len; };
    std::stable_sort(small.begin(), small.end(), by_len);
    std::stable_sort(large.begin(), large.end(), by_len);
    if (out_tiles) *out_tiles = (int64_t)(small.size() + large.size());

    const size_t inst_bytes = inst.size() * sizeof(InstDesc);
    const size_t small_bytes = small.size() * sizeof(Tile);
    const size_t large_bytes = large.size() * sizeof(Tile);
    const size_t zero_bytes = zero.size() * sizeof(int64_t);
    const size_t bytes = 256 + inst_bytes + small_bytes + large_bytes + zero_bytes;
    unsigned char* ws = static_cast<unsigned char*>(stream_workspace(s, bytes));
    if (!ws) return set_error(HYD_E_OOM, "batched_histogram: workspace allocation failed");
    unsigned int* counters = reinterpret_cast<unsigned int*>(ws);
    InstDesc* d_inst = reinterpret_cast<InstDesc*>(ws + 256);
    Tile* d_small = reinterpret_cast<Tile*>(ws + 256 + inst_bytes);
    Tile* d_large = reinterpret_cast<Tile*>(ws + 256 + inst_bytes + small_bytes);
    int64_t* d_zero = reinterpret_cast<int64_t*>(ws + 256 + inst_bytes + small_bytes + large_bytes);
    {
        // one staging buffer, one copy
        std::lock_guard<std::mutex> lk(g_plan_mu);
        std::vector<unsigned char> stage(bytes - 256);
        unsigned char* q = stage.data();
        memcpy(q, inst.data(), inst_bytes);
        memcpy(q + inst_bytes, small.data(), small_bytes);
        memcpy(q + inst_bytes + small_bytes, large.data(), large_bytes);
        memcpy(q + inst_bytes + small_bytes + large_bytes, zero.data(), zero_bytes);
        HYD_TRY_HIP(hipMemcpyAsync(ws + 256, stage.data(), stage.size(), hipMemcpyHostToDevice, s));
        HYD_TRY_HIP(hipStreamSynchronize(s));  // pageable source: consumed before it goes away
    }
    HYD_TRY(fill_zero(counters, 256, s));
    const DeviceInfo& di = device_info();
    if (!zero.empty()) {
        int64_t nr = (int64_t)zero.size() / 2;
        zero_ranges_kernel<<<(unsigned)std::min<int64_t>(nr, 4096), 256, 0, s>>>(d_zero, nr);
        HYD_CHECK_LAUNCH();
    }
    if (!small.empty()) {
        auto k = batched_hist_kernel<kSmallBlock, kSmallWords>;
        const int lds = kSmallWords * 8;
        HYD_TRY_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        int64_t blocks = std::min<int64_t>((int64_t)small.size(), (int64_t)di.num_cus * 4);
        k<<<(unsigned)blocks, kSmallBlock, lds, s>>>(d_inst, d_small, (int64_t)small.size(),
                                                     counters);
        HYD_CHECK_LAUNCH();
    }
    if (!large.empty()) {
        auto k = batched_hist_kernel<kLargeBlock, kLargeWords>;
        const int lds = kLargeWords * 8;
        HYD_TRY_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        int64_t blocks = std::min<int64_t>((int64_t)large.size(), (int64_t)di.num_cus);
        k<<<(unsigned)blocks, kLargeBlock, lds, s>>>(d_inst, d_large, (int64_t)large.size(),
                                                     counters + 1);
        HYD_CHECK_LAUNCH();
    }
    return HYD_OK;
}

// Packed layout: instance k reads data[off[k] ..] and writes hist[hoff[k] ..].
int batched_histogram(const int64_t* data, const int64_t* off, const int64_t* len,
                      const int64_t* bins, const int64_t* hoff, int64_t count, int64_t* hist,
                      int policy, hipStream_t s, int64_t* out_tiles) {
    std::vector<const int64_t*> src((size_t)count);
    std::vector<int64_t*> dst((size_t)count);
    for (int64_t k = 0; k < count; ++k) {
        if (off[k] < 0 || hoff[k] < 0)
            return set_error(HYD_E_INVALID, "batched_histogram: bad offsets at %lld", (long long)k);
        src[k] = data + off[k];
        dst[k] = hist + hoff[k];
    }
    return batched_histogram_ptrs(src.data(), len, bins, dst.data(), count, policy, s, out_tiles);
}

}  // namespace hyd

using namespace hyd;

extern "C" int hydhip_batched_histogram_i64(const int64_t* data, const int64_t* off,
                                            const int64_t* len, const int64_t* bins,
                                            const int64_t* hoff, int64_t count, int64_t* hist,
                                            int policy, void* stream) {
    HYD_TRY(ensure_init());
    if (count < 0 || (count > 0 && (!off || !len || !bins || !hoff || !hist)) || policy < 0 ||
        policy > 2)
        return set_error(HYD_E_INVALID, "hydhip_batched_histogram_i64: bad arguments");
    if (count == 0) return HYD_OK;
    return batched_histogram(data, off, len, bins, hoff, count, hist, policy, resolve_stream(stream),
                             nullptr);
}
