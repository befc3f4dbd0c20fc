// batch.hip - many uneven weighted-histogram instances in one persistent
// launch, with the privatization strategy picked per instance size class.
//
// Each instance k is the program of programs/weighted_histogram.hyd with its
// own bin count: hist_k[x % bins_k] += x*x over data_k (Parallelize.hs:466-503
// would privatize every one of them on the CPU).  On the GPU one launch
// serves the whole batch: the host cuts the instances into tiles (largest
// first), the persistent workgroups pull tiles from an atomic queue counter
// and run each tile either
//   * LDS mode:    private bins in LDS (several copies for small bin counts),
//                  one global atomic per touched bin at the end of the tile, or
//   * GLOBAL mode: one 64-bit global atomic per element (bins beyond LDS, or
//                  too few elements per bin to pay for zeroing and merging).
// The choice per (bin class, size class) comes from the autotuner table
// (tune.cpp) or, for classes it has not measured, a size rule.
// Integer + is order-free mod 2^64, so every schedule gives the oracle's bits.
#include <stdio.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "hyd_internal.hpp"

namespace hyd {

// Autotuner interface (tune.cpp).
int tuned_batch_mode(int64_t n, int64_t bins);  // -1 if the class is untuned

namespace {

constexpr int kBlock = 512;
constexpr int kLdsWords = 4096 + 1;           // LDS budget per workgroup (u64 words)
constexpr int64_t kMaxLdsBins = 4096;

struct Tile {
    int64_t inst;
    int64_t start;
    int64_t len;
    int64_t mode;  // 0 = LDS, 1 = global atomics
};

struct InstDesc {
    int64_t off;   // first element in data
    int64_t bins;
    int64_t hoff;  // first bin in hist
    int64_t pad;
};

__device__ __forceinline__ void lds_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Calls add(x) for every element of p[0..len) handled by this workgroup's
// threads (16-byte loads when aligned, 4 pairs in flight per thread).
template <class Add>
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
    for (; q + 3 * kBlock < npairs; q += 4 * kBlock) {
        longlong2 a = p2[q], b = p2[q + kBlock], c = p2[q + 2 * kBlock], d = p2[q + 3 * kBlock];
        add(a.x); add(a.y); add(b.x); add(b.y);
        add(c.x); add(c.y); add(d.x); add(d.y);
    }
    for (; q < npairs; q += kBlock) {
        longlong2 a = p2[q];
        add(a.x);
        add(a.y);
    }
    if ((len & 1) && threadIdx.x == 0) add(p[len - 1]);
}

__global__ __launch_bounds__(kBlock) void batched_hist_kernel(const int64_t* __restrict__ data,
                                                              const InstDesc* __restrict__ inst,
                                                              const Tile* __restrict__ tiles,
                                                              int64_t ntiles,
                                                              unsigned int* __restrict__ counter,
                                                              unsigned long long* __restrict__ hist) {
    __shared__ unsigned long long lds[kLdsWords];
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
        unsigned long long* h = hist + d.hoff;
        const int64_t* src = data + d.off + tile.start;
        if (tile.mode == 0) {
            const int stride = (int)(nb | 1);
            const int copies = max(1, min(16, kLdsWords / stride));
            const int words = copies * stride;
            for (int i = threadIdx.x; i < words; i += kBlock) lds[i] = 0;
            __syncthreads();
            unsigned long long* mine = lds + (threadIdx.x % copies) * stride;
            tile_elements(src, tile.len, [&](int64_t x) {
                int64_t b = x % nb;
                if ((uint64_t)b < (uint64_t)nb) lds_add(&mine[b], (uint64_t)x * (uint64_t)x);
            });
            __syncthreads();
            for (int b = threadIdx.x; b < nb; b += kBlock) {
                unsigned long long s = 0;
                for (int c = 0; c < copies; ++c) s += lds[c * stride + b];
                if (s) global_add(&h[b], s);
            }
            __syncthreads();  // LDS reused by the next tile
        } else {
            tile_elements(src, tile.len, [&](int64_t x) {
                int64_t b = x % nb;
                if ((uint64_t)b < (uint64_t)nb) global_add(&h[b], (uint64_t)x * (uint64_t)x);
            });
        }
    }
}

// Size rule for classes the autotuner has not measured: LDS when the bins fit
// and there are >= 8 elements per bin.
int rule_mode(int64_t n, int64_t bins) { return (bins <= kMaxLdsBins && n >= 8 * bins) ? 0 : 1; }

int64_t tile_len_for(int mode, int64_t bins) {
    if (mode == 0) return std::max<int64_t>(16384, 8 * bins);
    return 16384;
}

std::mutex g_plan_mu;

}  // namespace

// Host-side plan + launch.  Descriptor arrays are host memory; the plan is
// uploaded into the stream workspace together with the queue counter.
int batched_histogram(const int64_t* data, const int64_t* off, const int64_t* len,
                      const int64_t* bins, const int64_t* hoff, int64_t count, int64_t* hist,
                      int policy, hipStream_t s, int64_t* out_tiles) {
    std::vector<InstDesc> inst((size_t)count);
    std::vector<Tile> tiles;
    int64_t total_bins = 0;
    for (int64_t k = 0; k < count; ++k) {
        if (len[k] < 0 || bins[k] < 1 || off[k] < 0 || hoff[k] < 0)
            return set_error(HYD_E_INVALID, "batched_histogram: bad descriptor %lld", (long long)k);
        inst[k] = InstDesc{off[k], bins[k], hoff[k], 0};
        total_bins = std::max(total_bins, hoff[k] + bins[k]);
        int mode;
        if (policy == 1)
            mode = bins[k] <= kMaxLdsBins ? 0 : 1;
        else if (policy == 2)
            mode = 1;
        else {
            mode = tuned_batch_mode(len[k], bins[k]);
            if (mode < 0) mode = rule_mode(len[k], bins[k]);
            if (bins[k] > kMaxLdsBins) mode = 1;
        }
        int64_t tl = tile_len_for(mode, bins[k]);
        for (int64_t st = 0; st < len[k]; st += tl)
            tiles.push_back(Tile{k, st, std::min(tl, len[k] - st), mode});
    }
    // largest tiles first (longest-processing-time order for the queue)
    std::stable_sort(tiles.begin(), tiles.end(), [](const Tile& a, const Tile& b) {
        return a.len > b.len;
    });
    const int64_t ntiles = (int64_t)tiles.size();
    if (out_tiles) *out_tiles = ntiles;
    // hist: every instance's bins start at zero (fill [bins] 0)
    HYD_TRY(fill_zero(hist, (size_t)total_bins * 8, s));
    if (ntiles == 0) return HYD_OK;
    size_t inst_bytes = inst.size() * sizeof(InstDesc);
    size_t tile_bytes = tiles.size() * sizeof(Tile);
    size_t bytes = 256 + inst_bytes + tile_bytes;
    unsigned char* ws = static_cast<unsigned char*>(stream_workspace(s, bytes));
    if (!ws) return set_error(HYD_E_OOM, "batched_histogram: workspace allocation failed");
    unsigned int* counter = reinterpret_cast<unsigned int*>(ws);
    InstDesc* d_inst = reinterpret_cast<InstDesc*>(ws + 256);
    Tile* d_tiles = reinterpret_cast<Tile*>(ws + 256 + inst_bytes);
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        HYD_TRY_HIP(hipMemcpyAsync(d_inst, inst.data(), inst_bytes, hipMemcpyHostToDevice, s));
        HYD_TRY_HIP(hipMemcpyAsync(d_tiles, tiles.data(), tile_bytes, hipMemcpyHostToDevice, s));
        // pageable sources: make sure the copies have consumed the vectors
        HYD_TRY_HIP(hipStreamSynchronize(s));
    }
    HYD_TRY(fill_zero(counter, 256, s));
    const DeviceInfo& di = device_info();
    int64_t blocks = std::min<int64_t>(ntiles, (int64_t)di.num_cus * 4);
    batched_hist_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(
        data, d_inst, d_tiles, ntiles, counter, reinterpret_cast<unsigned long long*>(hist));
    HYD_CHECK_LAUNCH();
    return HYD_OK;
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
