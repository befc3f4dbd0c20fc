// sort_scan.hip - int64 scans, stable LSD radix sort of (key, index) pairs,
// gather, segmented reduce and the COO -> CSR kernels.
//
// Reference semantics:
//   scan family           src/Language/Hydrangea/Lowering.hs:3563-3607
//   segmented_reduce      Lowering.hs:874-938
//   sort_indices          Lowering.hs:961-965 -> runtime/hydrangea_runtime.h:391-456
//                         (stable LSD radix, 8-bit digits, passes = bytes of
//                         the largest key, at least one)
//   gather                Lowering.hs:1860-1898
//   coo_sum_duplicates    Lowering.hs:966-1067 (adjacent equal (row, col)
//                         merge; outputs sized like the input, zero tail)
//   csr_from_sorted_coo   Lowering.hs:1068-1147 (row counts, exclusive scan,
//                         row_ptr[nrows] = total, copy cols/vals)
//
// CDNA4 mapping: every scan is reduce-then-scan over 2048-element tiles in
// two launches (tile totals; then each tile adds up the totals before it
// itself and scans its elements), so no workgroup ever waits on another.
// Measured on MI355X (1e7 int64): 47 us, against 59-95 us for a single pass
// with decoupled look-back (a ticket per tile plus status round trips at
// device scope, which cross the XCDs' non-coherent L2s) and 54 us for three
// launches.  The keyed COO merge works the same way over run-head counts.
// The radix sort
// runs per 8-bit digit: per-tile digit counts (LDS atomics), the run starts
// from the count table (two launches over blocks of tiles), and a stable scatter that ranks each element
// inside its tile with wave64 ballots (a lane's rank among equal digits is
// the popcount of the equal-digit lanes below it) plus per-wave LDS counts.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "hyd_internal.hpp"

namespace hyd {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // scan tile: 2048 elements
// Up to this many tiles each tile adds up the totals before it itself (one
// pass over at most 8192 words of L2-resident totals); beyond it a third
// launch scans the totals first.
constexpr int64_t kRedundantTiles = 8192;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int64_t u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Exclusive scan of one value per thread across the workgroup; returns the
// exclusive prefix and writes the block total to *total (all threads).
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* total) {
    __shared__ int64_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t incl = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        int64_t s = wsum[w];
        if (w < wave) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

// Element e of the scanned sequence: forward a[e], reverse a[n-1-e].
__device__ __forceinline__ int64_t seq_index(int64_t e, int64_t n, bool reverse) {
    return reverse ? n - 1 - e : e;
}

// Elements of a scanned sequence.  ArraySrc: a[e] (forward) or a[n-1-e]
// (reverse).  HeadsSrc: 1 where a run of equal keys starts in a sorted key
// array, else 0 - the run-head flags, computed on the fly instead of written
// out and read back.
struct ArraySrc {
    const int64_t* a;
    int64_t n;
    bool reverse;
    __device__ __forceinline__ int64_t operator()(int64_t e) const { return a[seq_index(e, n, reverse)]; }
};
struct HeadsSrc {
    const int64_t* keys;
    __device__ __forceinline__ int64_t operator()(int64_t e) const {
        return (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
    }
};

template <class Src>
__global__ __launch_bounds__(kBlock) void scan_tile_sums(Src src, int64_t n, int64_t* __restrict__ sums) {
    const int64_t base = (int64_t)blockIdx.x * kTile;
    uint64_t v = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t e = base + (int64_t)k * kBlock + threadIdx.x;
        if (e < n) v += (uint64_t)src(e);
    }
    int64_t total;
    (void)block_excl_scan((int64_t)v, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the tile totals, in place, in chunks
// (OP 0: +, 1: max of non-negative values).
template <int OP>
__global__ __launch_bounds__(kBlock) void scan_sums_kernel(int64_t* __restrict__ sums, int64_t m) {
    __shared__ uint64_t wv[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (int64_t base = 0; base < m; base += kTile) {
        uint64_t v[kItems];
        uint64_t local = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int64_t e = base + (int64_t)threadIdx.x * kItems + k;
            v[k] = e < m ? (uint64_t)sums[e] : 0;
            local = OP == 0 ? local + v[k] : (v[k] > local ? v[k] : local);
        }
        uint64_t incl = local;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t o = (uint64_t)__shfl_up((long long)incl, off, 64);
            if (lane >= off) incl = OP == 0 ? incl + o : (o > incl ? o : incl);
        }
        if (lane == 63) wv[wave] = incl;
        __syncthreads();
        uint64_t pre = (uint64_t)__shfl_up((long long)incl, 1, 64);
        if (lane == 0) pre = 0;
        uint64_t all = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) {
            if (w < wave) pre = OP == 0 ? pre + wv[w] : (wv[w] > pre ? wv[w] : pre);
            all = OP == 0 ? all + wv[w] : (wv[w] > all ? wv[w] : all);
        }
        __syncthreads();
        pre = OP == 0 ? pre + carry : (carry > pre ? carry : pre);
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int64_t e = base + (int64_t)threadIdx.x * kItems + k;
            if (e < m) sums[e] = (int64_t)pre;
            pre = OP == 0 ? pre + v[k] : (v[k] > pre ? v[k] : pre);
        }
        carry = OP == 0 ? carry + all : (all > carry ? all : carry);
    }
}

// The combination (OP 0: +, 1: max) of the totals of the tiles before this
// workgroup's tile: read off a scanned table (scanned), or added up here
// from the raw totals (each thread a strided share, then the workgroup).
// Contains barriers: called by every thread.
template <int OP>
__device__ __forceinline__ uint64_t tiles_before(const int64_t* __restrict__ sums, int64_t tile, bool scanned) {
    __shared__ uint64_t wv[kBlock / 64];
    if (scanned) return (uint64_t)sums[tile];
    // eight totals in flight per thread at a time (one at a time, each
    // would wait out an L2 round trip)
    uint64_t v = 0;
    constexpr int U = 8;
    for (int64_t j0 = threadIdx.x; j0 < tile; j0 += (int64_t)U * kBlock) {
        uint64_t x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = j0 + (int64_t)u * kBlock;
            x[u] = j < tile ? (uint64_t)sums[j] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v = OP == 0 ? v + x[u] : (x[u] > v ? x[u] : v);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = (uint64_t)__shfl_xor((long long)v, off, 64);
        v = OP == 0 ? v + o : (o > v ? o : v);
    }
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t all = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) all = OP == 0 ? all + wv[w] : (wv[w] > all ? wv[w] : all);
    __syncthreads();
    return all;
}

// Per tile: out[e] = init + (exclusive or inclusive) prefix of the sequence,
// out indexed like the source (reversed for a reverse scan; out may alias
// the source: a tile reads all of its elements before it writes any).  The
// tile goes through LDS both ways: coalesced loads and stores (lane =
// consecutive element), while each thread scans kItems consecutive elements
// (row t of the tile, padded to an odd stride of 64-bit words against bank
// conflicts).  sums == NULL: a single tile.
constexpr int kRow = kItems + 1;

template <class Src>
__global__ __launch_bounds__(kBlock) void scan_apply(Src src, int64_t n, bool reverse, bool inclusive,
                                                     int64_t init, const int64_t* __restrict__ sums,
                                                     bool scanned, int64_t* __restrict__ out) {
    __shared__ int64_t tile[kBlock * kRow];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // element j = k * kBlock + t of the tile
        const int j = k * kBlock + t;
        const int64_t e = base + j;
        tile[(j / kItems) * kRow + j % kItems] = e < n ? src(e) : 0;
    }
    const uint64_t before_tile = sums ? tiles_before<0>(sums, blockIdx.x, scanned) : 0;
    __syncthreads();
    uint64_t v[kItems];
    uint64_t local = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = (uint64_t)tile[t * kRow + k];
        local += v[k];
    }
    int64_t total;
    uint64_t acc = (uint64_t)block_excl_scan((int64_t)local, &total) + before_tile + (uint64_t)init;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (inclusive) {
            acc += v[k];
            tile[t * kRow + k] = (int64_t)acc;
        } else {
            tile[t * kRow + k] = (int64_t)acc;
            acc += v[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int j = k * kBlock + t;
        const int64_t e = base + j;
        if (e < n) out[seq_index(e, n, reverse)] = tile[(j / kItems) * kRow + j % kItems];
    }
}

// Workspace of a scan: one word per tile (the totals).
template <class Src>
int scan_src(Src src, int64_t n, int64_t init, int64_t* out, bool inclusive, bool reverse, hipStream_t s,
             int64_t* tile_ws) {
    if (n == 0) return HYD_OK;
    const int64_t tiles = ceil_div(n, kTile);
    if (tiles > 0x7fffffff) return set_error(HYD_E_INVALID, "scan: too many elements");
    if (tiles == 1) {  // one launch
        scan_apply<Src><<<1, kBlock, 0, s>>>(src, n, reverse, inclusive, init, nullptr, false, out);
        HYD_CHECK_LAUNCH();
        return HYD_OK;
    }
    scan_tile_sums<Src><<<(unsigned)tiles, kBlock, 0, s>>>(src, n, tile_ws);
    HYD_CHECK_LAUNCH();
    const bool scanned = tiles > kRedundantTiles;
    if (scanned) {
        scan_sums_kernel<0><<<1, kBlock, 0, s>>>(tile_ws, tiles);
        HYD_CHECK_LAUNCH();
    }
    scan_apply<Src><<<(unsigned)tiles, kBlock, 0, s>>>(src, n, reverse, inclusive, init, tile_ws, scanned, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

}  // namespace

// Scan over device memory with a workspace from the stream (also used by the
// sort and the COO kernels).
int scan_i64(const int64_t* in, int64_t n, int64_t init, int64_t* out, bool inclusive, bool reverse,
             hipStream_t s, int64_t* tile_ws) {
    return scan_src(ArraySrc{in, n, reverse}, n, init, out, inclusive, reverse, s, tile_ws);
}

// out[i] = number of runs of equal keys that start at or before i in the
// sorted keys[0, n) (an inclusive scan of the run-head flags, which are
// never stored).  ws: scan_workspace_words(n).
int scan_heads_i64(const int64_t* keys, int64_t n, int64_t* out, hipStream_t s, int64_t* tile_ws) {
    return scan_src(HeadsSrc{keys}, n, 0, out, true, false, s, tile_ws);
}

int64_t scan_workspace_words(int64_t n) { return ceil_div(std::max<int64_t>(n, 1), kTile) + 1; }

namespace {

// ---- radix sort -----------------------------------------------------------
constexpr int kSortBlock = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortBlock * kSortRounds;  // 4096 keys per tile

__global__ __launch_bounds__(kBlock) void max_key_kernel(const int64_t* __restrict__ keys, int64_t n,
                                                         unsigned long long* __restrict__ out) {
    // max(0, max key) as unsigned: the pass count only needs non-negative
    // maxima (a negative maximum means one pass, like max_key = 0).  Eight
    // loads in flight per thread per trip; one device-scope atomic per
    // workgroup (atomics on one address serialise: one per wave of a
    // 1024-workgroup grid took 61 us for 1e7 keys).
    __shared__ int64_t wmax[kBlock / 64];
    int64_t m = 0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + 7 * stride < n; i += 8 * stride) {
        int64_t x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = __builtin_nontemporal_load(&keys[i + k * stride]);
#pragma unroll
        for (int k = 0; k < 8; ++k) m = x[k] > m ? x[k] : m;
    }
    for (; i < n; i += stride) m = keys[i] > m ? keys[i] : m;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        int64_t o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
        __hip_atomic_fetch_max(out, (unsigned long long)m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Digits of 8 or 9 bits: a pass covers `BITS` bits of the key at `shift`.
constexpr int kMaxDigitBits = 9;

template <int BITS>
__device__ __forceinline__ int digit_of(int64_t key, int shift) {
    return (int)(((uint64_t)key >> shift) & ((1u << BITS) - 1));
}

// counts[tile * 2^BITS + d] = number of keys of `tile` whose digit is d
// (tile-major: each tile writes its counts contiguously).
template <int BITS>
__global__ __launch_bounds__(kSortBlock) void radix_count_kernel(const int64_t* __restrict__ keys,
                                                                 int64_t n, int shift, int64_t ntiles,
                                                                 int64_t* __restrict__ counts) {
    constexpr int D = 1 << BITS;
    __shared__ unsigned int hist[4][D];
    for (int i = threadIdx.x; i < 4 * D; i += kSortBlock) (&hist[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    unsigned int* mine = hist[threadIdx.x >> 6];
    int64_t k[kSortRounds];  // all of the thread's loads in flight at once
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const int64_t e = base + (int64_t)r * kSortBlock + threadIdx.x;
        k[r] = e < n ? keys[e] : 0;
    }
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r)
        if (base + (int64_t)r * kSortBlock + threadIdx.x < n) atomicAdd(&mine[digit_of<BITS>(k[r], shift)], 1u);
    __syncthreads();
    for (int d = threadIdx.x; d < D; d += kSortBlock)
        counts[(int64_t)blockIdx.x * D + d] = (int64_t)hist[0][d] + hist[1][d] + hist[2][d] + hist[3][d];
}

// Run starts from the tile-major count table, in two launches over blocks of
// consecutive tiles (at most kSortBlocks blocks; thread i holds digits i,
// i + 256, ...: every access to a table row is coalesced):
//   radix_block_sums_kernel: bsum[b][d] = the block's count of digit d;
//   radix_offsets_kernel:    per block, the digit's start (the column totals,
//     then an exclusive scan over digits), plus the earlier blocks' counts
//     (each block adds those up itself), then down the block's rows:
//     offsets[t][d] = start[d] + counts of digit d in the tiles before t.
// `offsets` may be `counts` (each thread reads an entry before it writes it).
// No workgroup waits on another, and no table is transposed: measured on
// MI355X over 1e7 keys, 13.6 us per pass against 28 us for two transposes
// around a scan of the digit-major table.
constexpr int kSortBlocks = 64;

template <int BITS>
__global__ __launch_bounds__(kSortBlock) void radix_block_sums_kernel(const int64_t* __restrict__ counts,
                                                                      int64_t ntiles, int64_t per_block,
                                                                      int64_t* __restrict__ bsum) {
    constexpr int D = 1 << BITS, PER = D / kSortBlock;
    const int64_t t0 = (int64_t)blockIdx.x * per_block;
    const int64_t t1 = t0 + per_block < ntiles ? t0 + per_block : ntiles;
    int64_t acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = 0;
    constexpr int U = 16;  // rows in flight at a time
    for (int64_t tb = t0; tb < t1; tb += U) {
        int64_t v[U][PER];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < PER; ++q) v[u][q] = tb + u < t1 ? counts[(tb + u) * D + threadIdx.x + q * kSortBlock] : 0;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < PER; ++q) acc[q] += v[u][q];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) bsum[(int64_t)blockIdx.x * D + threadIdx.x + q * kSortBlock] = acc[q];
}

template <int BITS>
__global__ __launch_bounds__(kSortBlock) void radix_offsets_kernel(const int64_t* counts, int64_t ntiles,
                                                                   int64_t per_block, int64_t nblocks,
                                                                   const int64_t* __restrict__ bsum,
                                                                   int64_t* offsets) {
    constexpr int D = 1 << BITS, PER = D / kSortBlock;
    const int64_t b = blockIdx.x;
    int64_t total[PER], before[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) total[q] = before[q] = 0;
    // sixteen rows in flight at a time (a loop of dependent loads would wait
    // out one L2 round trip per block)
    constexpr int U = 16;
    for (int64_t c0 = 0; c0 < nblocks; c0 += U) {
        int64_t v[U][PER];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < PER; ++q)
                v[u][q] = c0 + u < nblocks ? bsum[(c0 + u) * D + threadIdx.x + q * kSortBlock] : 0;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                total[q] += v[u][q];
                before[q] += c0 + u < b ? v[u][q] : 0;
            }
    }
    // digit d's start: the totals of the digits below it (digits q * 256 + i
    // in order: the scan over threads for q = 0, then q = 1 after all of them)
    int64_t start[PER];
    int64_t carry = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int64_t all;
        const int64_t ex = block_excl_scan(total[q], &all);
        start[q] = carry + ex + before[q];
        carry += all;
    }
    const int64_t t0 = b * per_block;
    const int64_t t1 = t0 + per_block < ntiles ? t0 + per_block : ntiles;
    for (int64_t tb = t0; tb < t1; tb += U) {
        int64_t v[U][PER];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < PER; ++q)
                v[u][q] = tb + u < t1 ? counts[(tb + u) * D + threadIdx.x + q * kSortBlock] : 0;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (tb + u < t1) offsets[(tb + u) * D + threadIdx.x + q * kSortBlock] = start[q];
                start[q] += v[u][q];
            }
    }
}

// Stable scatter of one tile, keys with their payloads.  idx_in == nullptr
// means the identity (the first pass of an index sort).
//
// The tile's 4096 elements are split into four contiguous quarters, one per
// wave (element e = wave * 1024 + round * 64 + lane).  Each wave ranks its
// quarter alone, with no workgroup barrier: per round, the lanes holding the
// same digit find each other with wave64 ballots, read the wave's running
// count of that digit from LDS and the first of them adds the group's size
// (a wave's LDS operations take effect in program order).  One barrier later
// the slot of an element is the tile's start of its digit (an exclusive scan
// of the tile's digit counts, which the scanned count table holds as
// differences) + the digit's counts in earlier waves + its rank in its wave.
// The keys are then placed in LDS at their slots and written out in slot
// order - consecutive threads write consecutive addresses of a digit's run -
// and the payloads after them the same way.
//
// Four workgroups fit on a CU (4 waves per SIMD) with 8-bit digits: 40 KiB of LDS
// (the digit per slot in a byte, the wave sums kept in that array's
// space before it is written) and at most 128 VGPRs (the payloads are loaded
// after the ranking, while the keys are written out).  At three per CU the
// 2442 tiles of 1e7 keys took four rounds, the last one 18 % full.
template <int BITS>
__global__ __launch_bounds__(kSortBlock) __attribute__((amdgpu_waves_per_eu(BITS <= 8 ? 4 : 3)))
void radix_scatter_kernel(
    const int64_t* __restrict__ keys_in, const int64_t* __restrict__ idx_in, int64_t n, int shift,
    int64_t ntiles, const int64_t* __restrict__ offsets, int64_t* __restrict__ keys_out,
    int64_t* __restrict__ idx_out) {
    constexpr int D = 1 << BITS;
    constexpr int PER = D / kSortBlock;  // digits per thread in the tile scan
    constexpr int kWaves = kSortBlock / 64;
    constexpr int kQuarter = kSortTile / kWaves;  // elements per wave
    static_assert(PER >= 1 && D % kSortBlock == 0 && kQuarter == 64 * kSortRounds, "tile shape");
    using digit_t = typename std::conditional<(BITS <= 8), uint8_t, uint16_t>::type;
    __shared__ int64_t stage[kSortTile];       // keys, then payloads, in slot order
    __shared__ __attribute__((aligned(16))) digit_t slot_digit[kSortTile];  // the digit of each slot
    __shared__ int64_t gbase[D];               // global position of slot 0's digit run
    __shared__ uint16_t wave_run[kWaves][D];   // per wave: its count of each digit, then its start
    // the wave sums live in slot_digit's space: read before the barrier
    // that precedes the first write of slot_digit
    unsigned* wave_sum = reinterpret_cast<unsigned*>(slot_digit);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tile = blockIdx.x;
    const int64_t base = tile * kSortTile;
    const int tile_n = (int)(n - base < kSortTile ? n - base : kSortTile);

    for (int i = threadIdx.x; i < kWaves * D; i += kSortBlock) (&wave_run[0][0])[i] = 0;
    const int first = wave * kQuarter + lane;  // this lane's element of round 0
    int64_t key[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const int e = first + r * 64;
        key[r] = e < tile_n ? keys_in[base + e] : 0;
    }
    __syncthreads();

    // rank within the wave's quarter
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint16_t* mine = wave_run[wave];
    int rank[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const bool valid = first + r * 64 < tile_n;
        const int d = valid ? digit_of<BITS>(key[r], shift) : 0;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const uint64_t m = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const int below = __popcll(peers & lt_mask);
        int at = 0;
        if (valid) at = mine[d];
        rank[r] = at + below;
        if (valid && below == 0) mine[d] = (uint16_t)(at + __popcll(peers));
    }

    // the tile's digit counts (differences of the digit-major scanned table),
    // their exclusive scan, and each wave's start within each digit's run
    // `offsets` is tile-major: offsets[t * D + d] is where tile t's run of
    // digit d starts; the next start in digit-major order is tile t+1's, or
    // digit d+1's in tile 0 after the last tile
    unsigned cnt[PER], sum = 0;
    int64_t start[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int d = threadIdx.x * PER + q;
        start[q] = offsets[tile * D + d];
        const int64_t next = tile + 1 < ntiles ? offsets[(tile + 1) * D + d] : (d + 1 < D ? offsets[d + 1] : n);
        cnt[q] = (unsigned)(next - start[q]);
        sum += cnt[q];
    }
    unsigned incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();  // wave sums and every wave's digit counts are in
    unsigned excl = incl - sum;
    for (int w = 0; w < wave; ++w) excl += wave_sum[w];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int d = threadIdx.x * PER + q;
        gbase[d] = start[q] - excl;
        unsigned o = excl;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const unsigned c = wave_run[w][d];
            wave_run[w][d] = (uint16_t)o;
            o += c;
        }
        excl += cnt[q];
    }
    __syncthreads();

    // keys in slot order; the payloads' loads go out meanwhile
    int slot[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        slot[r] = 0;
        if (first + r * 64 < tile_n) {
            slot[r] = mine[digit_of<BITS>(key[r], shift)] + rank[r];
            stage[slot[r]] = key[r];
        }
    }
    int64_t idx[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const int e = first + r * 64;
        idx[r] = e < tile_n ? (idx_in ? idx_in[base + e] : base + e) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < tile_n; i += kSortBlock) {
        const int64_t k = stage[i];
        const int d = digit_of<BITS>(k, shift);
        slot_digit[i] = (uint16_t)d;
        keys_out[gbase[d] + i] = k;
    }
    __syncthreads();
    // payloads in slot order
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r)
        if (first + r * 64 < tile_n) stage[slot[r]] = idx[r];
    __syncthreads();
    for (int i = threadIdx.x; i < tile_n; i += kSortBlock) idx_out[gbase[slot_digit[i]] + i] = stage[i];
}

inline unsigned grid_for(int64_t n, int block = 256, int64_t cap = 8192) {
    int64_t g = ceil_div(n, block);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

}  // namespace

// Workspace layout for sort_pairs_i64: keys x2, payload x1 (the output is
// the other), digit counts 2^10 * ntiles twice (tile-major, digit-major),
// scan tile sums, 1 word for the max key.
int64_t sort_workspace_words(int64_t n) {
    int64_t ntiles = ceil_div(std::max<int64_t>(n, 1), kSortTile);
    const int64_t ncounts = ((int64_t)1 << kMaxDigitBits) * ntiles;
    return 2 * n + n + 2 * ncounts + scan_workspace_words(ncounts) + 8;
}

// Digit widths for sorting the low `bits` bits: 8-bit digits, or 9-bit ones
// where that saves a pass (the widths as even as possible; a pass may cover
// bits above `bits`, which are zero for keys that fit).  Measured on MI355X
// over 1e7 keys with payloads: a 10-bit pass 218 us, an 8-bit pass 133 us
// (256 runs per 4096-key tile write 16-element runs, 1024 runs write
// partial lines), so 40-bit keys take 5 passes of 8 bits, not 4 of 10.
static int digit_widths(int bits, int* widths) {
    bits = std::max(bits, 8);
    const int p8 = (bits + 7) / 8, p9 = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
    if (p9 < p8) {
        const int per = bits / p9, extra = bits % p9;
        for (int p = 0; p < p9; ++p) widths[p] = std::max(8, per + (p < extra ? 1 : 0));
        return p9;
    }
    for (int p = 0; p < p8; ++p) widths[p] = 8;
    return p8;
}

// One pass: per-tile digit counts (tile-major, so each tile writes one
// contiguous row), every (tile, digit) run's start from them in place (block
// sums, then offsets), and the scatter, each tile reading its starts as one
// contiguous row.  (Digit-major rows read and written directly cost one
// memory line per digit and tile: twice the pass's compulsory traffic at
// 10-bit digits.)
template <int BITS>
static int radix_pass(const int64_t* kin, const int64_t* iin, int64_t n, int shift, int64_t ntiles,
                      int64_t* counts, int64_t* dmajor, int64_t* kout, int64_t* iout,
                      hipStream_t s) {
    radix_count_kernel<BITS><<<(unsigned)ntiles, kSortBlock, 0, s>>>(kin, n, shift, ntiles, counts);
    HYD_CHECK_LAUNCH();
    const int64_t per_block = ceil_div(ntiles, kSortBlocks), nblocks = ceil_div(ntiles, per_block);
    int64_t* bsum = dmajor;  // nblocks * D words
    radix_block_sums_kernel<BITS><<<(unsigned)nblocks, kSortBlock, 0, s>>>(counts, ntiles, per_block, bsum);
    HYD_CHECK_LAUNCH();
    radix_offsets_kernel<BITS><<<(unsigned)nblocks, kSortBlock, 0, s>>>(counts, ntiles, per_block, nblocks, bsum,
                                                                        counts);
    HYD_CHECK_LAUNCH();
    radix_scatter_kernel<BITS><<<(unsigned)ntiles, kSortBlock, 0, s>>>(kin, iin, n, shift, ntiles, counts,
                                                                       kout, iout);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

// Stable LSD radix sort of keys[0..n) carrying a 64-bit payload per key
// (payload == nullptr: the key's index, which makes the output the sorting
// permutation).  Without a bound (max_key < 0) the sort orders the low
// 8 * B bits, B = the number of bytes of max(0, max key) and at least one,
// as the reference's byte passes do (negative keys included); the max is
// read back to the host once.  max_key >= 0 promises keys in [0, max_key]
// and sorts just its bits.  Each pass takes a digit of 8 or 9 bits.
// Sorted payloads go to payload_out, sorted keys to keys_out when it is
// non-null; *passes_out gets the reference's byte-pass count.  `ws` holds
// sort_workspace_words(n) words.
int sort_pairs_i64(const int64_t* keys, const int64_t* payload, int64_t n, int64_t max_key,
                   int64_t* keys_out, int64_t* payload_out, hipStream_t s, int64_t* ws, int* passes_out) {
    if (n == 0) return HYD_OK;
    int64_t ntiles = ceil_div(n, kSortTile);
    int64_t* kbuf0 = ws;
    int64_t* kbuf1 = ws + n;
    int64_t* ibuf = ws + 2 * n;
    int64_t* counts = ws + 3 * n;
    int64_t* dmajor = counts + ((int64_t)1 << kMaxDigitBits) * ntiles;
    int64_t* scan_ws = dmajor + ((int64_t)1 << kMaxDigitBits) * ntiles;
    int bits;
    if (max_key < 0) {
        unsigned long long* maxw = reinterpret_cast<unsigned long long*>(
            scan_ws + scan_workspace_words(((int64_t)1 << kMaxDigitBits) * ntiles));
        HYD_TRY(fill_zero(maxw, 8, s));
        max_key_kernel<<<grid_for(n, kBlock, device_info().num_cus), kBlock, 0, s>>>(keys, n, maxw);
        HYD_CHECK_LAUNCH();
        unsigned long long maxk = 0;
        HYD_TRY_HIP(hipMemcpyAsync(&maxk, maxw, 8, hipMemcpyDeviceToHost, s));
        HYD_TRY_HIP(hipStreamSynchronize(s));
        int bytes = 0;
        for (unsigned long long m = maxk; m; m >>= 8) ++bytes;
        bytes = std::max(bytes, 1);
        if (passes_out) *passes_out = bytes;
        bits = 8 * bytes;
    } else {
        bits = 64 - __builtin_clzll((unsigned long long)max_key | 1);
        if (passes_out) *passes_out = (bits + 7) / 8;
    }
    int widths[8];
    const int passes = digit_widths(bits, widths);

    // Ping-pong so that the last pass writes into payload_out (and keys_out).
    const int64_t* kin = keys;
    const int64_t* iin = payload;
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        int64_t* kout = (last && keys_out) ? keys_out : ((p & 1) ? kbuf1 : kbuf0);
        int64_t* iout = ((passes - 1 - p) & 1) ? ibuf : payload_out;
        if (widths[p] == 8)
            HYD_TRY(radix_pass<8>(kin, iin, n, shift, ntiles, counts, dmajor, kout, iout, s));
        else
            HYD_TRY(radix_pass<9>(kin, iin, n, shift, ntiles, counts, dmajor, kout, iout, s));
        shift += widths[p];
        kin = kout;
        iin = iout;
    }
    return HYD_OK;
}

// Stable sort of keys[0..n) -> perm (indices).
int sort_indices_i64(const int64_t* keys, int64_t n, int64_t* perm, hipStream_t s, int64_t* ws,
                     int* passes_out) {
    return sort_pairs_i64(keys, nullptr, n, -1, nullptr, perm, s, ws, passes_out);
}

namespace {

// ---- gather, segmented reduce, COO -> CSR ----------------------------------

__global__ void gather_kernel(const int64_t* __restrict__ perm, int64_t n,
                              const int64_t* __restrict__ src, int64_t nsrc,
                              int64_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = perm[i];
        out[i] = (uint64_t)p < (uint64_t)nsrc ? src[p] : 0;
    }
}

// One thread per segment for short segments; the reference's inner loop is a
// plain sequential reduction and int64 + is order-free mod 2^64.
__global__ void segmented_reduce_kernel(const int64_t* __restrict__ offsets, int64_t nseg,
                                        const int64_t* __restrict__ vals, int64_t nvals, int64_t init,
                                        int64_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t wave_id = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    // one wave per segment: lanes stride through the segment, then reduce
    for (int64_t sg = wave_id; sg < nseg; sg += nwaves) {
        int64_t start = offsets[sg], stop = offsets[sg + 1];
        uint64_t acc = 0;
        for (int64_t k = start + lane; k < stop; k += 64)
            acc += (uint64_t)((uint64_t)k < (uint64_t)nvals ? vals[k] : 0);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc += (uint64_t)__shfl_xor((long long)acc, off, 64);
        if (lane == 0) out[sg] = (int64_t)((uint64_t)init + acc);
    }
}

// keys[i] = rows[i] * ncols + cols[i]  (the packed row-major key of the COO
// benchmark program, int64 wraparound).
__global__ void pack_keys_kernel(const int64_t* __restrict__ rows, const int64_t* __restrict__ cols,
                                 int64_t n, int64_t ncols, int64_t* __restrict__ keys) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = (int64_t)((uint64_t)rows[i] * (uint64_t)ncols + (uint64_t)cols[i]);
}

// head[i] = 1 when element i starts a new (row, col) run of the (optionally
// permuted) sequence.
__global__ void coo_heads_kernel(const int64_t* __restrict__ perm, const int64_t* __restrict__ rows,
                                 const int64_t* __restrict__ cols, int64_t n,
                                 int64_t* __restrict__ head) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t pi = perm ? perm[i] : i;
        int64_t h = 1;
        if (i > 0) {
            int64_t pj = perm ? perm[i - 1] : i - 1;
            h = (rows[pi] != rows[pj] || cols[pi] != cols[pj]) ? 1 : 0;
        }
        head[i] = h;
    }
}

// Duplicates are merged without atomics: with P the inclusive prefix sum of
// the (permuted) values and slot(i) = inclusive count of heads - 1, the run
// ending at element i sums to P[i] - P[end of the previous run] (exact in
// wrapping int64 arithmetic, like the reference's run loop).  Heads write
// row/col into their slot; run ends record P into end_p[slot].
__global__ void coo_run_ends_kernel(const int64_t* __restrict__ perm, const int64_t* __restrict__ rows,
                                    const int64_t* __restrict__ cols, const int64_t* __restrict__ head,
                                    const int64_t* __restrict__ slot_incl, const int64_t* __restrict__ psum,
                                    int64_t n, int64_t* __restrict__ out_rows,
                                    int64_t* __restrict__ out_cols, int64_t* __restrict__ end_p) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t slot = slot_incl[i] - 1;
        if (head[i]) {
            const int64_t pi = perm ? perm[i] : i;
            out_rows[slot] = rows[pi];
            out_cols[slot] = cols[pi];
        }
        if (i == n - 1 || head[i + 1]) end_p[slot] = psum[i];
    }
}

// out_vals[k] = end_p[k] - end_p[k - 1] for the runs k < *count.
__global__ void coo_run_sums_kernel(const int64_t* __restrict__ end_p, const int64_t* __restrict__ count,
                                    int64_t n, int64_t* __restrict__ out_vals) {
    const int64_t cnt = *count;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n && k < cnt;
         k += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t prev = k ? (uint64_t)end_p[k - 1] : 0;
        out_vals[k] = (int64_t)((uint64_t)end_p[k] - prev);
    }
}

// ---- keyed COO -> CSR ---------------------------------------------------------
// Over (key, value) pairs sorted by key, keys one-to-one with (row, col)
// (row * ncols + col inside the box [0, nrows) x [0, ncols)):
//   scan_tile_sums<HeadsSrc>: the number of run heads in each tile.
//   coo_keyed_merge_kernel: element e's slot is the number of run heads at or
//     before it, minus one (the heads of earlier tiles added up by each tile,
//     as in the scans).  A head writes its column into col_idx[slot]; a run's
//     sum is the sum of its values (wrapping int64, as the reference's run
//     loop adds them), kept as a running sum that a segmented scan carries
//     across the threads of a tile.  The part of a run inside its head's tile
//     is stored in out_vals[slot]; a tile whose first run started in an
//     earlier tile records (slot, its part) in cont[], which the fix-up kernel
//     adds with one atomic per tile (integer + commutes, so the order does not
//     matter).  The last merged entry of each row records row_last1[row] =
//     slot + 1, and each row tile's largest such value goes into rmax[] (one
//     atomic max per row tile a merge tile ends rows in).
//   coo_rows_kernel: row_ptr[0] = 0, row_ptr[r + 1] = max(row_last1[0..r])
//     (rows are sorted, so that is the number of merged entries with row <= r:
//     the exclusive scan of the reference's row counts); the row tiles before
//     a tile contribute through rmax[].
//   coo_fix_kernel: the cont[] atomics and the zero tails past the count.
struct SegCarry {
    uint64_t sum;  // sum since the last run head (of everything when none)
    int head;      // 1 when the span holds a run head
};
__device__ __forceinline__ SegCarry seg_combine(SegCarry a, SegCarry b) {  // a before b
    return SegCarry{b.head ? b.sum : a.sum + b.sum, a.head | b.head};
}

template <bool POW2>
__device__ __forceinline__ int64_t key_row(uint64_t k, int col_shift, uint64_t ncols) {
    return POW2 ? (int64_t)(k >> col_shift) : (int64_t)(k / ncols);
}

template <bool POW2>
__global__ __launch_bounds__(kBlock) void coo_keyed_merge_kernel(
    const int64_t* __restrict__ skeys, const int64_t* __restrict__ svals, int64_t n, int64_t nrows,
    int64_t ncols, int col_shift, const int64_t* __restrict__ hsums, bool scanned,
    int64_t* __restrict__ col_idx, int64_t* __restrict__ out_vals, int64_t* __restrict__ row_last1,
    unsigned long long* __restrict__ rmax, int64_t* __restrict__ cont, int64_t* __restrict__ d_cnt) {
    __shared__ int64_t stage[kBlock * kRow];
    __shared__ int64_t edge[2];  // the key before the tile, the key after it
    __shared__ int64_t cont_slot_sh, cont_val_sh;
    __shared__ uint64_t wave_heads[kBlock / 64];
    __shared__ SegCarry wave_seg[kBlock / 64];
    const int64_t tno = blockIdx.x;
    const int64_t base = tno * kTile;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int tile_n = (int)(n - base < kTile ? n - base : kTile);
    const uint64_t ucols = (uint64_t)ncols;

    // keys and values loaded together (coalesced: lane = consecutive
    // element), then each moved through LDS to the thread that owns kItems
    // consecutive elements; the neighbours across the tile edges as well
    int64_t kin[kItems], vin[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int j = k * kBlock + t;
        kin[k] = j < tile_n ? skeys[base + j] : 0;
        vin[k] = j < tile_n ? svals[base + j] : 0;
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int j = k * kBlock + t;
        stage[(j / kItems) * kRow + j % kItems] = kin[k];
    }
    if (t == 0) {
        edge[0] = base > 0 ? skeys[base - 1] : 0;
        edge[1] = base + kTile < n ? skeys[base + kTile] : 0;
        cont_slot_sh = -1;
        cont_val_sh = 0;
    }
    // heads before the tile (the loads overlap the key loads)
    const uint64_t P = tiles_before<0>(hsums, tno, scanned);
    __syncthreads();
    uint64_t key[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) key[k] = (uint64_t)stage[t * kRow + k];
    const uint64_t key_prev = t > 0 ? (uint64_t)stage[(t - 1) * kRow + kItems - 1] : (uint64_t)edge[0];
    const uint64_t key_next = t + 1 < kBlock ? (uint64_t)stage[(t + 1) * kRow] : (uint64_t)edge[1];
    __syncthreads();
    // values, the same way
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int j = k * kBlock + t;
        stage[(j / kItems) * kRow + j % kItems] = vin[k];
    }
    __syncthreads();
    uint64_t val[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) val[k] = (uint64_t)stage[t * kRow + k];

    // run heads and ends of the thread's elements (bit k)
    const int64_t e0 = base + (int64_t)t * kItems;
    uint32_t head_bits = 0, end_bits = 0;
    uint64_t hc = 0;
    SegCarry mine{0, 0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t e = e0 + k;
        if (e < n) {
            const uint64_t prev = k > 0 ? key[k - 1] : key_prev;
            const uint64_t next = k + 1 < kItems ? key[k + 1] : key_next;
            const bool h = e == 0 || key[k] != prev;
            const bool en = e == n - 1 || key[k] != next;
            head_bits |= (uint32_t)h << k;
            end_bits |= (uint32_t)en << k;
            hc += h;
            mine = h ? SegCarry{val[k], 1} : SegCarry{mine.sum + val[k], mine.head};
        }
    }
    // exclusive scans over the threads: head counts and the segmented carry
    uint64_t hincl = hc;
    SegCarry sincl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t ho = (uint64_t)__shfl_up((long long)hincl, off, 64);
        SegCarry so;
        so.sum = (uint64_t)__shfl_up((long long)sincl.sum, off, 64);
        so.head = __shfl_up(sincl.head, off, 64);
        if (lane >= off) {
            hincl += ho;
            sincl = seg_combine(so, sincl);
        }
    }
    if (lane == 63) {
        wave_heads[wave] = hincl;
        wave_seg[wave] = sincl;
    }
    uint64_t hexcl = hincl - hc;
    SegCarry sexcl;
    sexcl.sum = (uint64_t)__shfl_up((long long)sincl.sum, 1, 64);
    sexcl.head = __shfl_up(sincl.head, 1, 64);
    if (lane == 0) sexcl = SegCarry{0, 0};
    __syncthreads();  // wave totals in; every thread has read its values from stage
    uint64_t tile_heads = 0;
    SegCarry wpre{0, 0};
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) {
            hexcl += wave_heads[w];
            wpre = seg_combine(wpre, wave_seg[w]);
        }
        tile_heads += wave_heads[w];
    }
    sexcl = seg_combine(wpre, sexcl);
    if (t == 0 && tno == (int64_t)gridDim.x - 1) *d_cnt = (int64_t)(P + tile_heads);

    // The tile's heads own the slots P .. P + tile_heads - 1.  Outputs go
    // through LDS indexed by slot - P and are stored in slot order, so that
    // consecutive lanes write consecutive addresses (a thread's own slots are
    // kItems elements apart from its neighbours').  Two rounds over the one
    // LDS tile: columns, then run sums.
    const int th = (int)tile_heads;
    const int last_t = (tile_n - 1) / kItems, last_k = (tile_n - 1) % kItems;
    int lslot = (int)hexcl - 1;  // local slot of the run in progress
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (e0 + k >= n) break;
        if ((head_bits >> k) & 1) {
            ++lslot;
            stage[lslot] = POW2 ? (int64_t)(key[k] & (ucols - 1)) : (int64_t)(key[k] % ucols);
        }
    }
    __syncthreads();
    for (int i = t; i < th; i += kBlock) col_idx[P + i] = stage[i];
    __syncthreads();
    lslot = (int)hexcl - 1;
    uint64_t running = sexcl.sum;
    bool headed = sexcl.head != 0;  // the run in progress has its head in this tile
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t e = e0 + k;
        if (e >= n) break;
        if ((head_bits >> k) & 1) {
            ++lslot;
            running = val[k];
            headed = true;
        } else {
            running += val[k];
        }
        const bool en = (end_bits >> k) & 1;
        const bool tile_last = t == last_t && k == last_k;
        if (en || tile_last) {
            // the run's part inside this tile: the whole run when it has its
            // head here (or its first part, when it goes on into the next
            // tile), else the continuation of a run from an earlier tile
            if (headed) {
                stage[lslot] = (int64_t)running;
            } else {
                cont_slot_sh = (int64_t)P - 1;
                cont_val_sh = (int64_t)running;
            }
        }
        if (en) {
            const uint64_t next = k + 1 < kItems ? key[k + 1] : key_next;
            const int64_t row = key_row<POW2>(key[k], col_shift, ucols);
            const bool last_of_all = e == n - 1;
            const int64_t nrow = last_of_all ? row : key_row<POW2>(next, col_shift, ucols);
            if ((last_of_all || nrow != row) && (uint64_t)row < (uint64_t)nrows) {
                const int64_t v = (int64_t)P + lslot + 1;
                row_last1[row] = v;
                // the last row end of its row tile within this merge tile
                if (last_of_all || tile_last || nrow / kTile != row / kTile)
                    __hip_atomic_fetch_max(&rmax[row / kTile], (unsigned long long)v, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    for (int i = t; i < th; i += kBlock) out_vals[P + i] = stage[i];
    if (t == 0) {
        cont[2 * tno] = cont_slot_sh;
        cont[2 * tno + 1] = cont_val_sh;
    }
}

// row_ptr[0] = 0, row_ptr[r + 1] = max(last1[0..r]) for r < nrows (values are
// counts, so max over non-negative words); the row tiles before a tile
// contribute their maxima rmax[] (scanned: already the exclusive max-scan).
__global__ __launch_bounds__(kBlock) void coo_rows_kernel(const int64_t* __restrict__ last1, int64_t nrows,
                                                          const int64_t* __restrict__ rmax, bool scanned,
                                                          int64_t* __restrict__ row_ptr) {
    __shared__ int64_t tile[kBlock * kRow];
    __shared__ uint64_t wmax[kBlock / 64];
    const int64_t tno = blockIdx.x;
    const int64_t base = tno * kTile;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int j = k * kBlock + t;
        const int64_t r = base + j;
        tile[(j / kItems) * kRow + j % kItems] = r < nrows ? last1[r] : 0;
    }
    const uint64_t before = tiles_before<1>(rmax, tno, scanned);
    __syncthreads();
    uint64_t v[kItems];
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = (uint64_t)tile[t * kRow + k];
        m = v[k] > m ? v[k] : m;
    }
    uint64_t incl = m;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t o = (uint64_t)__shfl_up((long long)incl, off, 64);
        if (lane >= off) incl = o > incl ? o : incl;
    }
    if (lane == 63) wmax[wave] = incl;
    uint64_t excl = (uint64_t)__shfl_up((long long)incl, 1, 64);
    if (lane == 0) excl = 0;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w)
        if (w < wave) excl = wmax[w] > excl ? wmax[w] : excl;
    uint64_t acc = before > excl ? before : excl;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        acc = v[k] > acc ? v[k] : acc;
        tile[t * kRow + k] = (int64_t)acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int j = k * kBlock + t;
        const int64_t r = base + j;
        if (r < nrows) row_ptr[r + 1] = tile[(j / kItems) * kRow + j % kItems];
    }
    if (tno == 0 && t == 0) row_ptr[0] = 0;
}

// The continuation atomics (one per tile whose first run started earlier)
// and the zero tails col_idx / out_vals [count, len).
__global__ void coo_fix_kernel(const int64_t* __restrict__ cont, int64_t ntiles,
                               const int64_t* __restrict__ d_cnt, int64_t len, int64_t* __restrict__ col_idx,
                               int64_t* __restrict__ out_vals) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = i0; i < ntiles; i += stride) {
        const int64_t slot = cont[2 * i];
        if (slot >= 0)
            __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(&out_vals[slot]),
                                   (unsigned long long)cont[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int64_t cnt = *d_cnt;
    for (int64_t i = cnt + i0; i < len; i += stride) {
        col_idx[i] = 0;
        out_vals[i] = 0;
    }
}

// keys[i] = rows[i] * ncols + cols[i]; *bad = 1 when any element lies
// outside [0, nrows) x [0, ncols) (one device-scope atomic per workgroup at
// most).  bad zeroed by the caller.  Flat form: one pair of elements per
// thread (16-byte accesses when rows, cols and keys are 16-byte aligned; the
// grid sized by the data, as the streaming map kernels): rows and cols are
// read once (non-temporal), the keys stay cached for the sort that reads
// them next.
__device__ __forceinline__ bool pack_one(int64_t r, int64_t c, int64_t nrows, int64_t ncols, int64_t* key) {
    *key = (int64_t)((uint64_t)r * (uint64_t)ncols + (uint64_t)c);
    return !(r >= 0 && r < nrows && c >= 0 && c < ncols);
}

__global__ __launch_bounds__(256) void pack_keys_box_kernel(const int64_t* __restrict__ rows,
                                                            const int64_t* __restrict__ cols, int64_t n,
                                                            int64_t nrows, int64_t ncols,
                                                            int64_t* __restrict__ keys,
                                                            unsigned long long* __restrict__ bad) {
    bool out = false;
    const bool vec = ((reinterpret_cast<uintptr_t>(rows) | reinterpret_cast<uintptr_t>(cols) |
                       reinterpret_cast<uintptr_t>(keys)) & 15) == 0;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // pair q: elements 2q, 2q + 1
    if (vec) {
        if (2 * q + 1 < n) {
            const v2i r = __builtin_nontemporal_load(reinterpret_cast<const v2i*>(rows) + q);
            const v2i c = __builtin_nontemporal_load(reinterpret_cast<const v2i*>(cols) + q);
            int64_t a, b;
            out |= pack_one(r.x, c.x, nrows, ncols, &a);
            out |= pack_one(r.y, c.y, nrows, ncols, &b);
            reinterpret_cast<v2i*>(keys)[q] = v2i{a, b};
        } else if (2 * q < n) {
            out |= pack_one(rows[2 * q], cols[2 * q], nrows, ncols, &keys[2 * q]);
        }
    } else {
        for (int64_t j = 2 * q; j < 2 * q + 2 && j < n; ++j) out |= pack_one(rows[j], cols[j], nrows, ncols, &keys[j]);
    }
    __shared__ int wbad[4];
    const bool any_out = __any(out);
    if ((threadIdx.x & 63) == 0) wbad[threadIdx.x >> 6] = any_out ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0 && (wbad[0] | wbad[1] | wbad[2] | wbad[3]))
        __hip_atomic_fetch_or(bad, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// flag = 1 when rows[0..n) is not non-decreasing (benign race: all write 1).
__global__ void unsorted_flag_kernel(const int64_t* __restrict__ rows, int64_t n, int64_t* flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (rows[i - 1] > rows[i]) *flag = 1;
}

// Sorted rows (flag 0): the first and the last element of each row's run
// record first[r] = i and last1[r] = i + 1, so the row's count is
// last1[r] - first[r] (0 - 0 for a row with no elements).  No atomics: each
// entry has exactly one writer.
__global__ void row_bounds_kernel(const int64_t* __restrict__ rows, int64_t nnz, int64_t nrows,
                                  const int64_t* __restrict__ flag, int64_t* __restrict__ first,
                                  int64_t* __restrict__ last1) {
    if (*flag) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = rows[i];
        if ((uint64_t)r >= (uint64_t)nrows) continue;
        if (i == 0 || rows[i - 1] != r) first[r] = i;
        if (i == nnz - 1 || rows[i + 1] != r) last1[r] = i + 1;
    }
}

__global__ void row_count_from_bounds_kernel(const int64_t* __restrict__ first,
                                             const int64_t* __restrict__ last1, int64_t nrows,
                                             const int64_t* __restrict__ flag, int64_t* __restrict__ counts) {
    if (*flag) return;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
         r += (int64_t)gridDim.x * blockDim.x)
        counts[r] = last1[r] - first[r];
}

// Rows in any order (flag 1): one global atomic per element.
__global__ void row_count_kernel(const int64_t* __restrict__ rows, int64_t nnz, int64_t nrows,
                                 const int64_t* __restrict__ flag, int64_t* __restrict__ counts) {
    if (!*flag) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = rows[i];
        if ((uint64_t)r < (uint64_t)nrows)
            __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(&counts[r]), 1ull,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// out[i] = i < nnz ? src[i] : 0 for i < len  (CSR copies cols/vals and keeps
// the zero tail of the freshly allocated output).
__global__ void copy_prefix_kernel(const int64_t* __restrict__ src, int64_t nnz, int64_t len,
                                   int64_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = i < nnz ? src[i] : 0;
}

// ---- order-exact fp64 scatter (+.) ------------------------------------------
// fp64 + is not associative, so the bits of a scatter (+.) depend on the order
// in which each bin receives its values: the reference's loop goes in element
// order (Lowering.hs:1432-1517, serial `for i`).  The GPU reproduces that
// order: a stable sort of the contributions by bin keeps element order within
// a bin, and the first element of each bin's run folds the run in that order.

// keys[i] = the bin of element i, or ndst when it does not contribute (a
// false guard or a route outside [0, ndst)): such elements sort last.
__global__ void scatter_keys_kernel(const int64_t* __restrict__ idx, const int64_t* __restrict__ guard,
                                    int64_t n, int64_t ndst, int64_t* __restrict__ keys) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = idx[i];
        keys[i] = ((!guard || guard[i]) && b >= 0 && b < ndst) ? b : ndst;
    }
}

// ---- order-preserving compaction of the contributing elements ------------
// Element i contributes when its guard holds (guard NULL: always) and its
// route idx[i] is a bin in [0, ndst).  Two passes over tiles of 16 rounds of
// 256 elements, element e = tile * 4096 + round * 256 + thread:
//  compact_count_kernel  one 64-bit ballot per (round, wave) and the tile's
//                        count;
//  compact_write_kernel  after a scan of the tile counts, writes the
//                        contributing keys and value bits in element order:
//                        rank = earlier rounds + earlier waves of this round
//                        + set bits of the ballot below the lane.
// The second pass reads the ballots instead of the guard, and idx and vals
// only where an element contributes.
constexpr int kCompactRounds = 16;
constexpr int64_t kCompactTile = 256 * kCompactRounds;

__global__ __launch_bounds__(256) void compact_count_kernel(const int64_t* __restrict__ idx,
                                                            const int64_t* __restrict__ guard, int64_t n,
                                                            int64_t ndst, uint64_t* __restrict__ masks,
                                                            int64_t* __restrict__ counts) {
    __shared__ int wave_total[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kCompactTile + threadIdx.x;
    bool keep[kCompactRounds];
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        const int64_t e = base + r * 256;
        keep[r] = false;
        if (e < n) {
            const int64_t bin = idx[e];
            keep[r] = (!guard || guard[e] != 0) && bin >= 0 && bin < ndst;
        }
    }
    int total = 0;
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        const uint64_t m = __ballot(keep[r]);
        if (lane == 0) masks[(int64_t)blockIdx.x * (kCompactRounds * 4) + r * 4 + wave] = m;
        total += __popcll(m);
    }
    if (lane == 0) wave_total[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0)
        counts[blockIdx.x] = (int64_t)wave_total[0] + wave_total[1] + wave_total[2] + wave_total[3];
}

__global__ __launch_bounds__(256) void compact_write_kernel(const int64_t* __restrict__ idx,
                                                            const int64_t* __restrict__ vals, int64_t n,
                                                            const uint64_t* __restrict__ masks,
                                                            const int64_t* __restrict__ incl_counts,
                                                            int64_t* __restrict__ ckeys,
                                                            int64_t* __restrict__ cvals) {
    __shared__ uint64_t tile_masks[kCompactRounds * 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < kCompactRounds * 4)
        tile_masks[threadIdx.x] = masks[(int64_t)blockIdx.x * (kCompactRounds * 4) + threadIdx.x];
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * kCompactTile + threadIdx.x;
    int64_t out = blockIdx.x ? incl_counts[blockIdx.x - 1] : 0;
#pragma unroll 4
    for (int r = 0; r < kCompactRounds; ++r) {
        const uint64_t* rm = tile_masks + r * 4;
        const uint64_t m = rm[wave];
        if ((m >> lane) & 1) {
            int64_t pos = out + __popcll(m & below);
            for (int w = 0; w < wave; ++w) pos += __popcll(rm[w]);
            const int64_t e = base + r * 256;  // < n: set bits only come from e < n
            ckeys[pos] = idx[e];
            cvals[pos] = vals[e];
        }
        out += __popcll(rm[0]) + __popcll(rm[1]) + __popcll(rm[2]) + __popcll(rm[3]);
    }
    (void)n;
}

// With r = (inclusive count of heads up to j) - 1: the first element of run r
// records its start, the last its end.
__global__ void run_bounds_kernel(const int64_t* __restrict__ skeys, const int64_t* __restrict__ incl,
                                  int64_t n, int64_t ndst, int64_t* __restrict__ start,
                                  int64_t* __restrict__ end) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = skeys[j];
        if (b >= ndst) continue;
        const int64_t r = incl[j] - 1;
        if (j == 0 || skeys[j - 1] != b) start[r] = j;
        if (j == n - 1 || skeys[j + 1] != b) end[r] = j + 1;
    }
}

// Folds the runs: old = dst[b]; new = val + old for the run's values in order
// (the combine write of Lowering.hs:3553-3557, operands in the reference's
// order), the values already in sorted order.  Short runs: one thread each
// (fold_short_runs_kernel).  A long run (more than kLongRun values - a grid
// whose points hit few voxels) gets a whole wave (fold_long_runs_kernel).
constexpr int64_t kLongRun = 32;

__global__ void fold_short_runs_kernel(const int64_t* __restrict__ skeys, const double* __restrict__ svals,
                                       const int64_t* __restrict__ start, const int64_t* __restrict__ end,
                                       const int64_t* __restrict__ nruns, double* __restrict__ dst) {
    const int64_t R = *nruns;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = start[r], e = end[r];
        if (e - a > kLongRun) continue;
        const int64_t b = skeys[a];
        double acc = dst[b];
        for (int64_t t = a; t < e; ++t) {
            const double v = svals[t];
            acc = v + acc;
        }
        dst[b] = acc;
    }
}

__device__ __forceinline__ double lane_value(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Exponent of the lowest set bit of a finite binary64 magnitude u (|x| =
// M * 2^k with M odd), INT_MAX for zero.  Subnormals hold mant * 2^-1074,
// normals (2^52 | mant) * 2^(E-1075); bit 63 stands in for "no bit" so that
// the count is defined for zero.  Straight-line code, no branches.
__device__ __forceinline__ int low_bit_exponent(uint64_t u) {
    const int E = (int)(u >> 52);
    const uint64_t M = (u & 0xFFFFFFFFFFFFFull) | ((uint64_t)(E != 0) << 52) | (1ull << 63);
    const int k = (int)__builtin_ctzll(M) + (E > 1 ? E : 1) - 1075;
    return u != 0 ? k : INT_MAX;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x = x + __shfl_xor(x, off, 64);
    return x;
}

// One wave per run index; waves whose run is short leave at once.
//
// Exact shortcut: let every operand (the old value and the run's values) be
// finite, 2^k the largest power of two dividing all of them and S the sum of
// their magnitudes.  If S < 2^(53+k), every partial sum in ANY order is an
// integer multiple of 2^k below 2^(53+k) in magnitude, hence representable:
// no addition rounds, and the serial fold and a tree sum both give the exact
// sum.  Zero results agree too: under round-to-nearest an exact zero sum is
// -0 only when both operands are -0, so either order yields -0 iff every
// operand is -0 (the lane sums start from -0, the identity that keeps it).
// S is summed in floating point in some order; with all terms multiples of
// 2^k the computed S reaches 2^(53+k) iff the true S does (partial sums are
// exact below that bound, rounding is monotone and the bound is
// representable), so the test is exact.  Dyadic data (weights n/64, integer
// counts) takes the shortcut; anything else folds serially: the 64 lanes
// load a step of 64 values together, the next step's loads in flight, and
// the adds run in order on the broadcast values.
__global__ __launch_bounds__(256) void fold_long_runs_kernel(const int64_t* __restrict__ skeys,
                                                             const double* __restrict__ svals,
                                                             const int64_t* __restrict__ start,
                                                             const int64_t* __restrict__ end,
                                                             const int64_t* __restrict__ nruns,
                                                             double* __restrict__ dst) {
    const int64_t R = *nruns;
    const int lane = threadIdx.x & 63;
    // wave-uniform indices: scalar loads of the bounds, scalar loop control
    const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int waves = (int)(gridDim.x * (blockDim.x >> 6));
    for (int64_t r = wid; r < R; r += waves) {
        const int64_t a = start[r], e = end[r];
        if (e - a <= kLongRun) continue;
        const int64_t b = skeys[a];
        const double old = dst[b];
        int kmin = INT_MAX;
        bool finite = true;
        double sabs = -0.0, part = -0.0;
        auto take = [&](double x) {
            const uint64_t u = __builtin_bit_cast(uint64_t, x) & 0x7FFFFFFFFFFFFFFFull;
            finite &= (u >> 52) != 0x7FF;
            kmin = min(kmin, low_bit_exponent(u));
            sabs = __builtin_bit_cast(double, u) + sabs;
            part = x + part;
        };
        if (lane == 0) take(old);
        // 16 loads in flight per lane; the tail's loads are clamped into the run
        int64_t t = a + lane;
        for (; t + 64 * 15 < e; t += 64 * 16) {
            double x[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) x[j] = svals[t + 64 * j];
#pragma unroll
            for (int j = 0; j < 16; ++j) take(x[j]);
        }
        {
            double x[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) x[j] = svals[min(t + 64 * j, e - 1)];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (t + 64 * j < e) take(x[j]);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) kmin = min(kmin, __shfl_xor(kmin, off, 64));
        const bool all_finite = __all(finite);
        sabs = wave_sum(sabs);
        part = wave_sum(part);
        bool exact = all_finite;
        if (exact && kmin != INT_MAX)  // kmin == INT_MAX: every operand is a zero
            exact = kmin <= 970 &&
                    sabs < __builtin_bit_cast(double, (uint64_t)(53 + kmin + 1023) << 52);
        // lane 0 decides for the wave (its S may round differently; any order is valid)
        if (__builtin_amdgcn_readfirstlane((int)exact)) {
            if (lane == 0) dst[b] = part;
            continue;
        }
        // Serial fold.  Loads are unconditional (clamped to the run's last
        // value) so that the wait before a step's adds covers only that step.
        double acc = old;
        double v = svals[min(a + lane, e - 1)];
        for (int64_t t0 = a; t0 < e; t0 += 64) {
            const double vn = svals[min(t0 + 64 + lane, e - 1)];  // next step in flight
            if (e - t0 >= 64) {
#pragma unroll
                for (int i = 0; i < 64; ++i) acc = lane_value(v, i) + acc;
            } else {
                const int cnt = (int)(e - t0);
                for (int i = 0; i < cnt; ++i) acc = lane_value(v, i) + acc;
            }
            v = vn;
        }
        if (lane == 0) dst[b] = acc;
    }
}

__global__ void set_word_kernel(int64_t* __restrict__ dst, const int64_t* __restrict__ a,
                                const int64_t* __restrict__ b) {
    // *dst = *a + *b  (e.g. row_ptr[nrows] = last prefix + last count)
    *dst = *a + *b;
}

}  // namespace

int gather_i64(const int64_t* perm, int64_t n, const int64_t* src, int64_t nsrc, int64_t* out,
               hipStream_t s) {
    if (n == 0) return HYD_OK;
    gather_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, src, nsrc, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

// coo_sum_duplicates over the (optionally permuted) sequence.  Outputs are
// `len` long with a zero tail; the deduplicated count goes to *d_nnz_out
// (device word).  ws: coo_workspace_words(n) words.
int coo_sum_duplicates_i64(const int64_t* perm, const int64_t* rows, const int64_t* cols,
                           const int64_t* vals, int64_t n, int64_t len, int64_t* out_rows,
                           int64_t* out_cols, int64_t* out_vals, int64_t* d_nnz_out, hipStream_t s,
                           int64_t* ws) {
    HYD_TRY(fill_zero(out_rows, (size_t)len * 8, s));
    HYD_TRY(fill_zero(out_cols, (size_t)len * 8, s));
    HYD_TRY(fill_zero(out_vals, (size_t)len * 8, s));
    if (n == 0) return fill_zero(d_nnz_out, 8, s);
    int64_t* head = ws;
    int64_t* incl = ws + n;
    int64_t* gv = ws + 2 * n;    // values in sequence order
    int64_t* psum = ws + 3 * n;  // their inclusive prefix sums
    int64_t* end_p = ws + 4 * n;
    int64_t* scan_ws = ws + 5 * n;
    coo_heads_kernel<<<grid_for(n), 256, 0, s>>>(perm, rows, cols, n, head);
    HYD_CHECK_LAUNCH();
    HYD_TRY(scan_i64(head, n, 0, incl, true, false, s, scan_ws));
    const int64_t* seq_vals = vals;
    if (perm) {
        gather_kernel<<<grid_for(n), 256, 0, s>>>(perm, n, vals, n, gv);
        HYD_CHECK_LAUNCH();
        seq_vals = gv;
    }
    HYD_TRY(scan_i64(seq_vals, n, 0, psum, true, false, s, scan_ws));
    coo_run_ends_kernel<<<grid_for(n), 256, 0, s>>>(perm, rows, cols, head, incl, psum, n, out_rows,
                                                    out_cols, end_p);
    HYD_CHECK_LAUNCH();
    coo_run_sums_kernel<<<grid_for(n), 256, 0, s>>>(end_p, incl + (n - 1), n, out_vals);
    HYD_CHECK_LAUNCH();
    HYD_TRY_HIP(hipMemcpyAsync(d_nnz_out, incl + (n - 1), 8, hipMemcpyDeviceToDevice, s));
    return HYD_OK;
}

int64_t coo_workspace_words(int64_t n) { return 5 * n + scan_workspace_words(n) + 8; }

// csr_from_sorted_coo with nnz known on the host.  Row counts without atomics
// when the rows are sorted (the usual case: the COO pipeline sorts by
// row-major key), one atomic per element otherwise; a device flag picks the
// path, so there is no host round trip.  ws: csr_workspace_words(nrows).
int csr_from_sorted_coo_i64(int64_t nrows, int64_t nnz, const int64_t* rows, const int64_t* cols,
                            const int64_t* vals, int64_t len, int64_t* row_ptr, int64_t* col_idx,
                            int64_t* out_vals, hipStream_t s, int64_t* ws) {
    int64_t* counts = ws;                       // nrows + 1
    int64_t* first = ws + (nrows + 1);          // nrows
    int64_t* last1 = ws + 2 * nrows + 1;        // nrows
    int64_t* flag = ws + 3 * nrows + 1;         // 1
    int64_t* scan_ws = ws + 3 * nrows + 2;
    HYD_TRY(fill_zero(ws, (size_t)(3 * nrows + 2) * 8, s));
    if (nnz > 0 && nrows > 0) {
        unsorted_flag_kernel<<<grid_for(nnz), 256, 0, s>>>(rows, nnz, flag);
        HYD_CHECK_LAUNCH();
        row_bounds_kernel<<<grid_for(nnz), 256, 0, s>>>(rows, nnz, nrows, flag, first, last1);
        HYD_CHECK_LAUNCH();
        row_count_from_bounds_kernel<<<grid_for(nrows), 256, 0, s>>>(first, last1, nrows, flag, counts);
        HYD_CHECK_LAUNCH();
        row_count_kernel<<<grid_for(nnz), 256, 0, s>>>(rows, nnz, nrows, flag, counts);
        HYD_CHECK_LAUNCH();
    }
    // exclusive scan over nrows+1 entries (the extra zero count makes
    // row_ptr[nrows] the total)
    HYD_TRY(scan_i64(counts, nrows + 1, 0, row_ptr, false, false, s, scan_ws));
    // a column or value array that already is the output (written there by
    // the duplicate merge, zero tail included) needs no copy
    if (len > 0 && cols != col_idx) {
        copy_prefix_kernel<<<grid_for(len), 256, 0, s>>>(cols, nnz, len, col_idx);
        HYD_CHECK_LAUNCH();
    }
    if (len > 0 && vals != out_vals) {
        copy_prefix_kernel<<<grid_for(len), 256, 0, s>>>(vals, nnz, len, out_vals);
        HYD_CHECK_LAUNCH();
    }
    return HYD_OK;
}

int64_t csr_workspace_words(int64_t nrows) { return 3 * nrows + 2 + scan_workspace_words(nrows + 1) + 8; }

// Order-exact scatter (+.) of vals (n, element order) routed by idx into
// dst[0, ndst): element i contributes when guard[i] != 0 (guard NULL: every
// element) and 0 <= idx[i] < ndst.  ws: scatter_ordered_workspace_words(n).
int scatter_ordered_add_f64(const int64_t* idx, const int64_t* guard, const double* vals, int64_t n,
                            int64_t ndst, double* dst, hipStream_t s, int64_t* ws) {
    if (n == 0) return HYD_OK;
    const int64_t ntiles = ceil_div(n, kCompactTile);
    int64_t* incl = ws + n;        // m: inclusive count of run heads
    int64_t* ckeys = ws + 2 * n;   // m: the contributing keys, element order
    int64_t* cvals = ws + 3 * n;   // m: their values' bits
    int64_t* skeys = ws + 5 * n;   // m: sorted keys
    int64_t* svals = ws + 6 * n;   // m: the values' bits in sorted order
    int64_t* start = ws + 7 * n;   // runs
    int64_t* end = ws + 8 * n;     // runs
    uint64_t* masks = reinterpret_cast<uint64_t*>(ws + 9 * n);  // 64 per tile
    int64_t* tcounts = ws + 9 * n + 64 * ntiles;                // per tile, then scanned
    int64_t* scratch = tcounts + ntiles;                         // sort, then scans
    // only the contributing elements are sorted (the guard often drops most)
    compact_count_kernel<<<(unsigned)ntiles, 256, 0, s>>>(idx, guard, n, ndst, masks, tcounts);
    HYD_CHECK_LAUNCH();
    HYD_TRY(scan_i64(tcounts, ntiles, 0, tcounts, true, false, s, scratch));
    compact_write_kernel<<<(unsigned)ntiles, 256, 0, s>>>(idx, reinterpret_cast<const int64_t*>(vals), n, masks,
                                                         tcounts, ckeys, cvals);
    HYD_CHECK_LAUNCH();
    int64_t m = 0;
    HYD_TRY_HIP(hipMemcpyAsync(&m, tcounts + (ntiles - 1), 8, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipStreamSynchronize(s));
    if (m == 0) return HYD_OK;
    // contributing keys lie in [0, ndst): that bound sets the pass count
    HYD_TRY(sort_pairs_i64(ckeys, cvals, m, ndst - 1, skeys, svals, s, scratch, nullptr));
    HYD_TRY(scan_heads_i64(skeys, m, incl, s, scratch));  // run index + 1 of each element
    run_bounds_kernel<<<grid_for(m), 256, 0, s>>>(skeys, incl, m, ndst, start, end);
    HYD_CHECK_LAUNCH();
    const double* sv = reinterpret_cast<const double*>(svals);
    fold_short_runs_kernel<<<grid_for(m), 256, 0, s>>>(skeys, sv, start, end, incl + (m - 1), dst);
    HYD_CHECK_LAUNCH();
    fold_long_runs_kernel<<<grid_for(m * 64 / kLongRun), 256, 0, s>>>(skeys, sv, start, end, incl + (m - 1), dst);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

int64_t scatter_ordered_workspace_words(int64_t n) {
    const int64_t ntiles = ceil_div(std::max<int64_t>(n, 1), kCompactTile);
    return 9 * n + 65 * ntiles + std::max(sort_workspace_words(n), scan_workspace_words(n)) + 8;
}

int pack_keys_i64(const int64_t* rows, const int64_t* cols, int64_t n, int64_t ncols, int64_t* keys,
                  hipStream_t s) {
    if (n == 0) return HYD_OK;
    pack_keys_kernel<<<grid_for(n), 256, 0, s>>>(rows, cols, n, ncols, keys);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

int segmented_reduce_i64(const int64_t* offsets, int64_t nseg, const int64_t* vals, int64_t nvals,
                         int64_t init, int64_t* out, hipStream_t s) {
    if (nseg == 0) return HYD_OK;
    segmented_reduce_kernel<<<grid_for(nseg * 64), 256, 0, s>>>(offsets, nseg, vals, nvals, init, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

}  // namespace hyd

namespace hyd {
int64_t coo_to_csr_workspace_words(int64_t nnz, int64_t nrows);
int coo_to_csr_ws(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* rows, const int64_t* cols,
                  const int64_t* vals, int64_t* row_ptr, int64_t* col_idx, int64_t* out_vals,
                  int64_t* nnz_out, hipStream_t s, int64_t* ws);
}  // namespace hyd

using namespace hyd;

extern "C" int hydhip_scan_i64(const int64_t* in, int64_t n, int64_t init, int64_t* out,
                               int inclusive, int reverse, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || (n > 0 && (!in || !out))) return set_error(HYD_E_INVALID, "hydhip_scan_i64: bad arguments");
    if (n == 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)scan_workspace_words(n) * 8));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_scan_i64: workspace allocation failed");
    return scan_i64(in, n, init, out, inclusive != 0, reverse != 0, s, ws);
}

extern "C" int hydhip_sort_indices_i64(const int64_t* keys, int64_t n, int64_t* perm, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || (n > 0 && (!keys || !perm)))
        return set_error(HYD_E_INVALID, "hydhip_sort_indices_i64: bad arguments");
    if (n == 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)sort_workspace_words(n) * 8));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_sort_indices_i64: workspace allocation failed");
    return sort_indices_i64(keys, n, perm, s, ws, nullptr);
}

extern "C" int hydhip_gather_i64(const int64_t* perm, int64_t n, const int64_t* src, int64_t nsrc,
                                 int64_t* out, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || nsrc < 0 || (n > 0 && (!perm || !out || (nsrc > 0 && !src))))
        return set_error(HYD_E_INVALID, "hydhip_gather_i64: bad arguments");
    return gather_i64(perm, n, src, nsrc, out, resolve_stream(stream));
}

extern "C" int hydhip_segmented_reduce_i64(const int64_t* offsets, int64_t noffsets,
                                           const int64_t* vals, int64_t nvals, int64_t init,
                                           int64_t* out, void* stream) {
    HYD_TRY(ensure_init());
    if (noffsets < 1 || nvals < 0 || !offsets || (noffsets > 1 && !out) || (nvals > 0 && !vals))
        return set_error(HYD_E_INVALID, "hydhip_segmented_reduce_i64: bad arguments");
    return segmented_reduce_i64(offsets, noffsets - 1, vals, nvals, init, out, resolve_stream(stream));
}

extern "C" int hydhip_coo_sum_duplicates_i64(const int64_t* rows, const int64_t* cols,
                                             const int64_t* vals, int64_t nnz, int64_t len,
                                             int64_t* out_rows, int64_t* out_cols,
                                             int64_t* out_vals, int64_t* nnz_out, void* stream) {
    HYD_TRY(ensure_init());
    if (nnz < 0 || len < nnz || !out_rows || !out_cols || !out_vals || !nnz_out ||
        (nnz > 0 && (!rows || !cols || !vals)))
        return set_error(HYD_E_INVALID, "hydhip_coo_sum_duplicates_i64: bad arguments");
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)coo_workspace_words(nnz) * 8 + 64));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_coo_sum_duplicates_i64: workspace allocation failed");
    int64_t* d_cnt = ws + coo_workspace_words(nnz);
    HYD_TRY(coo_sum_duplicates_i64(nullptr, rows, cols, vals, nnz, len, out_rows, out_cols, out_vals,
                                   d_cnt, s, ws));
    HYD_TRY_HIP(hipMemcpyAsync(nnz_out, d_cnt, 8, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipStreamSynchronize(s));
    return HYD_OK;
}

extern "C" int hydhip_csr_from_sorted_coo_i64(int64_t nrows, int64_t nnz, const int64_t* rows,
                                              const int64_t* cols, const int64_t* vals, int64_t len,
                                              int64_t* row_ptr, int64_t* col_idx, int64_t* out_vals,
                                              void* stream) {
    HYD_TRY(ensure_init());
    if (nrows < 0 || nnz < 0 || len < nnz || !row_ptr || (len > 0 && (!col_idx || !out_vals)) ||
        (nnz > 0 && (!rows || !cols || !vals)))
        return set_error(HYD_E_INVALID, "hydhip_csr_from_sorted_coo_i64: bad arguments");
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)csr_workspace_words(nrows) * 8));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_csr_from_sorted_coo_i64: workspace allocation failed");
    return csr_from_sorted_coo_i64(nrows, nnz, rows, cols, vals, len, row_ptr, col_idx, out_vals, s, ws);
}

// The whole COO -> CSR program of bench/coo_csr_build/coo_csr_build.hyd:
// packed keys, stable sort_indices, gather (fused into the duplicate merge),
// coo_sum_duplicates, csr_from_sorted_coo.  Outputs: row_ptr[nrows+1],
// col_idx[nnz], vals[nnz] (zero tails past *nnz_out), canonical rows/cols
// (out_rows, out_cols, may be NULL -> workspace).
extern "C" int hydhip_coo_to_csr_i64(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* rows,
                                     const int64_t* cols, const int64_t* vals, int64_t* row_ptr,
                                     int64_t* col_idx, int64_t* out_vals, int64_t* nnz_out,
                                     void* stream) {
    HYD_TRY(ensure_init());
    if (nrows < 0 || ncols < 0 || nnz < 0 || !row_ptr || !nnz_out ||
        (nnz > 0 && (!rows || !cols || !vals || !col_idx || !out_vals)))
        return set_error(HYD_E_INVALID, "hydhip_coo_to_csr_i64: bad arguments");
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(
        stream_workspace(s, (size_t)coo_to_csr_workspace_words(nnz, nrows) * 8));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_coo_to_csr_i64: workspace allocation failed");
    return coo_to_csr_ws(nrows, ncols, nnz, rows, cols, vals, row_ptr, col_idx, out_vals, nnz_out, s, ws);
}

namespace hyd {
// Words of workspace coo_to_csr_ws needs: keys, perm (or sorted keys),
// sorted values, canonical rows/cols/vals, and the largest of the sort /
// merge / CSR scratch, then 2 words of range-check results.
// Words of the keyed path's own region, after the general layout: 8 words
// (count, out-of-box flag), the row tiles' maxima and row_last1[nrows] (all
// zeroed per call), then the tiles' head counts and the merge's continuation
// records (2 per tile).
static int64_t coo_keyed_words(int64_t nnz, int64_t nrows) {
    const int64_t tm = ceil_div(std::max<int64_t>(nnz, 1), kTile), tr = ceil_div(std::max<int64_t>(nrows, 1), kTile);
    return 8 + tr + nrows + 3 * tm;
}

int64_t coo_to_csr_workspace_words(int64_t nnz, int64_t nrows) {
    int64_t scratch = std::max(sort_workspace_words(nnz),
                               std::max(coo_workspace_words(nnz), csr_workspace_words(nrows)));
    return 6 * nnz + scratch + 8 + coo_keyed_words(nnz, nrows);
}

// Inside the box [0, nrows) x [0, ncols) (and nrows * ncols < 2^62) the packed
// key is one-to-one with (row, col): sort (key, value) pairs by the box's key
// bits, then the keyed merge and the row scan above, with one read-back of
// (count, out-of-box flag).  When the flag shows an element
// outside the box, the result is recomputed by the reference's sequence:
// sort_indices, the merge comparing the (row, col) pairs themselves (colliding
// keys kept apart), csr_from_sorted_coo.
int coo_to_csr_ws(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* rows, const int64_t* cols,
                  const int64_t* vals, int64_t* row_ptr, int64_t* col_idx, int64_t* out_vals,
                  int64_t* nnz_out, hipStream_t s, int64_t* ws) {
    const int64_t keyed_words = coo_keyed_words(nnz, nrows);
    int64_t scratch = coo_to_csr_workspace_words(nnz, nrows) - 6 * nnz - 8 - keyed_words;
    int64_t* keys = ws;
    int64_t* perm = ws + nnz;  // or the sorted keys
    int64_t* crow = ws + 2 * nnz;
    int64_t* ccol = ws + 3 * nnz;
    int64_t* cval = ws + 4 * nnz;
    int64_t* svals = ws + 5 * nnz;
    int64_t* scr = ws + 6 * nnz;
    int64_t* d_cnt = scr + scratch;
    const bool box_fits = nrows > 0 && ncols > 0 && (unsigned __int128)nrows * (unsigned __int128)ncols <
                                                         ((unsigned __int128)1 << 62);
    if (nnz > 0 && box_fits) {
        const int64_t tm = ceil_div(nnz, kTile), tr = ceil_div(nrows, kTile);
        if (tm > 0x7fffffff || tr > 0x7fffffff || ceil_div(ceil_div(nnz, 2), 256) > 0x7fffffff)
            return set_error(HYD_E_INVALID, "coo_to_csr: too many elements");
        int64_t* X = d_cnt + 8;
        int64_t* k_cnt = X + 2;
        unsigned long long* bad = reinterpret_cast<unsigned long long*>(X + 3);
        unsigned long long* rmax = reinterpret_cast<unsigned long long*>(X + 8);
        int64_t* row_last1 = X + 8 + tr;
        int64_t* hsums = row_last1 + nrows;
        int64_t* cont = hsums + tm;
        HYD_TRY(fill_zero(X, (size_t)(8 + tr + nrows) * 8, s));
        pack_keys_box_kernel<<<(unsigned)ceil_div(ceil_div(nnz, 2), 256), 256, 0, s>>>(rows, cols, nnz, nrows,
                                                                                    ncols, keys, bad);
        HYD_CHECK_LAUNCH();
        // keys inside the box are at most nrows * ncols - 1: sorting those bits
        // orders them as the reference's byte passes do (extra high passes of
        // non-negative keys move nothing)
        HYD_TRY(sort_pairs_i64(keys, vals, nnz, nrows * ncols - 1, perm, svals, s, scr, nullptr));
        scan_tile_sums<HeadsSrc><<<(unsigned)tm, kBlock, 0, s>>>(HeadsSrc{perm}, nnz, hsums);
        HYD_CHECK_LAUNCH();
        const bool scanned_m = tm > kRedundantTiles, scanned_r = tr > kRedundantTiles;
        if (scanned_m) {
            scan_sums_kernel<0><<<1, kBlock, 0, s>>>(hsums, tm);
            HYD_CHECK_LAUNCH();
        }
        const int col_shift = (ncols & (ncols - 1)) == 0 ? __builtin_ctzll((uint64_t)ncols) : -1;
        if (col_shift >= 0)
            coo_keyed_merge_kernel<true><<<(unsigned)tm, kBlock, 0, s>>>(perm, svals, nnz, nrows, ncols, col_shift,
                                                                         hsums, scanned_m, col_idx, out_vals,
                                                                         row_last1, rmax, cont, k_cnt);
        else
            coo_keyed_merge_kernel<false><<<(unsigned)tm, kBlock, 0, s>>>(perm, svals, nnz, nrows, ncols, 0, hsums,
                                                                          scanned_m, col_idx, out_vals, row_last1,
                                                                          rmax, cont, k_cnt);
        HYD_CHECK_LAUNCH();
        coo_fix_kernel<<<grid_for(std::max(tm, nnz)), 256, 0, s>>>(cont, tm, k_cnt, nnz, col_idx, out_vals);
        HYD_CHECK_LAUNCH();
        if (scanned_r) {
            scan_sums_kernel<1><<<1, kBlock, 0, s>>>(reinterpret_cast<int64_t*>(rmax), tr);
            HYD_CHECK_LAUNCH();
        }
        coo_rows_kernel<<<(unsigned)tr, kBlock, 0, s>>>(row_last1, nrows, reinterpret_cast<const int64_t*>(rmax),
                                                        scanned_r, row_ptr);
        HYD_CHECK_LAUNCH();
        int64_t back[2];
        HYD_TRY_HIP(hipMemcpyAsync(back, k_cnt, sizeof back, hipMemcpyDeviceToHost, s));
        HYD_TRY_HIP(hipStreamSynchronize(s));
        if (!back[1]) {
            *nnz_out = back[0];
            return HYD_OK;
        }
        // an element outside the box: the general sequence below
    }
    HYD_TRY(pack_keys_i64(rows, cols, nnz, ncols, keys, s));
    HYD_TRY(sort_indices_i64(keys, nnz, perm, s, scr, nullptr));
    HYD_TRY(coo_sum_duplicates_i64(perm, rows, cols, vals, nnz, nnz, crow, ccol, cval, d_cnt, s, scr));
    int64_t cnt = 0;
    HYD_TRY_HIP(hipMemcpyAsync(&cnt, d_cnt, 8, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipStreamSynchronize(s));
    *nnz_out = cnt;
    return csr_from_sorted_coo_i64(nrows, cnt, crow, ccol, cval, nnz, row_ptr, col_idx, out_vals, s, scr);
}
}  // namespace hyd

// Guarded scatter (+.) on fp64, bit-exact to the reference's serial loop:
// dst (holding the defaults) receives vals[i] at idx[i] where guard[i] != 0
// (guard NULL: every element), bin by bin in element order.
extern "C" int hydhip_scatter_guarded_add_f64(const int64_t* idx, const double* vals,
                                              const int64_t* guard, int64_t n, double* dst,
                                              int64_t ndst, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || ndst < 1 || ndst > (int64_t)1 << 31 || !dst || (n > 0 && (!idx || !vals)))
        return set_error(HYD_E_INVALID, "hydhip_scatter_guarded_add_f64: bad arguments");
    if (n == 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)scatter_ordered_workspace_words(n) * 8));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_scatter_guarded_add_f64: workspace allocation failed");
    return scatter_ordered_add_f64(idx, guard, vals, n, ndst, dst, s, ws);
}

// Scatter chain (EScatterChain, Lowering.hs:1622-1858) with (+.) on fp64: the
// phases' contributions concatenated in phase order, then one order-exact
// scatter - per bin, phase 0's values in element order, then phase 1's, ...
// as the reference's phase loops produce.  idx/vals/guard/lens are host
// arrays of device pointers and lengths (guard[k] NULL: unguarded phase).
extern "C" int hydhip_scatter_chain_add_f64(int64_t nphases, const int64_t* const* idx,
                                            const double* const* vals, const int64_t* const* guard,
                                            const int64_t* lens, double* dst, int64_t ndst, void* stream) {
    HYD_TRY(ensure_init());
    if (nphases < 0 || ndst < 1 || ndst > (int64_t)1 << 31 || !dst ||
        (nphases > 0 && (!idx || !vals || !lens)))
        return set_error(HYD_E_INVALID, "hydhip_scatter_chain_add_f64: bad arguments");
    int64_t total = 0;
    for (int64_t k = 0; k < nphases; ++k) {
        if (lens[k] < 0 || (lens[k] > 0 && (!idx[k] || !vals[k])))
            return set_error(HYD_E_INVALID, "hydhip_scatter_chain_add_f64: bad phase %lld", (long long)k);
        total += lens[k];
    }
    if (total == 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    int64_t* ws = static_cast<int64_t*>(
        stream_workspace(s, (size_t)(2 * total + scatter_ordered_workspace_words(total)) * 8));
    if (!ws) return set_error(HYD_E_OOM, "hydhip_scatter_chain_add_f64: workspace allocation failed");
    int64_t* keys = ws;
    double* cvals = reinterpret_cast<double*>(ws + total);
    int64_t off = 0;
    for (int64_t k = 0; k < nphases; ++k) {
        const int64_t m = lens[k];
        if (m == 0) continue;
        scatter_keys_kernel<<<grid_for(m), 256, 0, s>>>(idx[k], guard ? guard[k] : nullptr, m, ndst, keys + off);
        HYD_CHECK_LAUNCH();
        HYD_TRY_HIP(hipMemcpyAsync(cvals + off, vals[k], (size_t)m * 8, hipMemcpyDeviceToDevice, s));
        off += m;
    }
    return scatter_ordered_add_f64(keys, nullptr, cvals, total, ndst, dst, s, ws + 2 * total);
}
