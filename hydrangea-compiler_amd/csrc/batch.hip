// batch.hip - many uneven weighted-histogram instances in one persistent
// launch, with the privatization strategy picked per instance size class.
//
// Each instance k is the program of programs/weighted_histogram.hyd with its
// own bin count: hist_k[x % bins_k] += x*x over data_k (Parallelize.hs:466-503
// privatizes every one of them on the CPU).  On the GPU the host cuts the
// batch into tiles and one persistent launch (one 1024-thread workgroup per
// CU, 128 KiB of LDS each) works through them.  A tile is (instance, element
// range, bin range, mode):
//   * OWNER:   the workgroup holds the tile's bin range in LDS (several copies
//              for small bin counts), streams every element of the instance and
//              writes the bin range with plain stores - no atomics, no zeroing.
//              Instances with more bins than LDS holds get one tile per
//              16384-bin range; each of them streams the instance.
//   * MERGE:   OWNER over one piece of an instance much longer than a
//              workgroup's share of the batch (one workgroup streams at a
//              fraction of HBM's rate): the piece's bins are added into the
//              pre-zeroed output with one global atomic per nonzero bin.
//   * GLOBAL:  one 64-bit global atomic per element of an element chunk into
//              the pre-zeroed output (the autotuner, tune.hip, can pick it per
//              (bin class, size class); global atomics on gfx950 run at tens of
//              G/s, so OWNER is the default).
// Scheduling: every workgroup has its own queue, planned on the host
// longest-first (LPT).  Measured on MI355X, device-scope atomics on one
// address serialise at close to a microsecond each, so a shared queue counter
// costs more than the tiles; per-workgroup counters are contention-free, and
// each workgroup claims its next tile while it works on the current one.
// Workgroup i of the launch starts on XCD i mod 8 (the dispatcher deals
// workgroups round-robin over the XCDs): the range tiles of one instance, which
// all stream the same elements, go to queues of workgroups on one XCD, so they
// run side by side and share its L2 - the instance comes from HBM about once
// instead of once per range.  A workgroup whose queue is empty takes tiles from
// the other queues (its own XCD first) and leaves when all are empty.  A wrong
// guess of the placement costs locality, never correctness.
// Integer + is order-free mod 2^64, so every schedule gives the oracle's bits.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <queue>
#include <vector>

#include "hyd_internal.hpp"

namespace hyd {

// Autotuner interface (tune.hip): 0 = LDS tiles, 1 = global atomics, -1 untuned.
// A snapshot of the whole table (one lock per batch, not per instance).
void tuned_batch_snapshot(TuneSnapshot* out);
int tune_class_bins(int64_t bins);
int tune_class_n(int64_t n);

namespace {

enum TileMode : int32_t { kModeOwner = 0, kModeMerge = 1, kModeGlobal = 2 };

// Self-contained: one load gives a workgroup everything about its tile.
struct Tile {
    const int64_t* src;       // first element of the tile
    unsigned long long* dst;  // the instance's bins
    int64_t len;              // elements
    int64_t bins;
    ModConst mod;             // route = x % bins
    int32_t lo;               // first bin of the range
    int32_t width;            // bins in the range
    int32_t mode;
    int32_t once;             // the tile is the only reader of its elements
};

constexpr int kBlock = 1024;
constexpr int kWords = 16384;                 // LDS words per workgroup: one bin range
constexpr int kXcds = 8;                      // XCDs of the MI355X
constexpr int64_t kChunk = (int64_t)1 << 18;  // elements per GLOBAL tile
constexpr int64_t kTileOverhead = 4096;       // fixed cost of a tile, in elements
constexpr int64_t kMinPiece = (int64_t)1 << 16;  // elements: the shortest piece of a split instance
constexpr int kPiecesPerQueue = 4;            // split so that each queue gets about this many pieces
constexpr int kTeam = 4;                      // workgroups that stream one instance together

__device__ __forceinline__ void lds_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_add(unsigned long long* p, uint64_t v) {
    __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Element pointers come out of the instance descriptors, so the compiler
// cannot tell they are global memory; without the cast it emits flat loads,
// which share a wait counter with the LDS atomics and serialise the loop.
typedef const __attribute__((address_space(1))) v2i GlobalPair;
typedef const __attribute__((address_space(1))) int64_t GlobalElem;

// Calls add(x) for every element of p[0..len) handled by this workgroup's
// threads (16-byte loads when aligned, 4 pairs in flight per thread).
// ONCE: the elements are read by this tile only - non-temporal loads (measured
// on MI355X: plain loads of a read-once stream keep the CU's L1 allocating and
// stall the address path for about half the cycles).  Otherwise plain loads,
// so that the other range tiles of the instance find the elements in L2.
template <int BLOCK, bool ONCE, class Add>
__device__ __forceinline__ void tile_elements(const int64_t* p_generic, int64_t len, const int64_t* spare, Add add) {
    if (len <= 0) return;
    GlobalElem* p = (GlobalElem*)p_generic;
    if (reinterpret_cast<uintptr_t>(p_generic) & 15) {
        if (threadIdx.x == 0) add(p[0]);
        ++p;
        --len;
    }
    GlobalPair* p2 = (GlobalPair*)p;
    const int64_t npairs = len >> 1;
    int64_t q = threadIdx.x;
    // software pipelined over two register sets (no copies between them, so
    // the wait before the arithmetic on one set leaves the other's loads in
    // flight): the next four pairs load while the current eight elements are
    // routed and added
    v2i A[4], B[4];
    // `spare`: 16 bytes the workgroup has just read (so in L2), the address
    // of the set loaded after the last one - it keeps the load count static
    // without fetching anything from HBM (a read-once stream's lines are not
    // kept in L2, so reloading the current set's addresses would fetch them
    // again: measured, 8.5% more HBM reads)
    GlobalPair* dummy = (GlobalPair*)spare;
    auto load = [&](v2i* r, bool real, int64_t at) {
        GlobalPair* base = real ? p2 + at : dummy;
        const int64_t step = real ? BLOCK : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (ONCE)
                r[k] = __builtin_nontemporal_load(&base[k * step]);
            else
                r[k] = base[k * step];
        }
    };
    auto run = [&](const v2i* r) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            add(r[k].x);
            add(r[k].y);
        }
    };
    // The next set is loaded unconditionally (after the last set from
    // `spare`; the values are then unused) so that the wait before the
    // arithmetic counts only the current set.
    if (q + 3 * BLOCK < npairs) {
        load(A, true, q);
        for (;;) {
            const bool nb = q + 7 * BLOCK < npairs;  // B = the four pairs after A
            load(B, nb, q + 4 * BLOCK);
            __builtin_amdgcn_sched_barrier(0);
            run(A);
            q += 4 * BLOCK;
            if (!nb) break;
            const bool na = q + 7 * BLOCK < npairs;  // A = the four pairs after B
            load(A, na, q + 4 * BLOCK);
            __builtin_amdgcn_sched_barrier(0);
            run(B);
            q += 4 * BLOCK;
            if (!na) break;
        }
    }
    for (; q < npairs; q += BLOCK) {
        const v2i a = p2[q];
        add(a.x);
        add(a.y);
    }
    if ((len & 1) && threadIdx.x == 0) add(p[len - 1]);
}

// Claims the next tile of queue v: its index, or -1 when the queue is empty.
__device__ __forceinline__ int64_t claim(const int64_t* __restrict__ qseg, unsigned int* counters, int v) {
    const int64_t b = qseg[v], e = qseg[v + 1];
    if (b >= e) return -1;
    const int64_t i = (int64_t)atomicAdd(&counters[v], 1u);
    return b + i < e ? b + i : -1;
}

// Run by the 64 lanes of wave 0 once the workgroup's own queue is empty: the
// lanes check up to 64 queues at a time for tiles left (own XCD first, then
// all; each workgroup starts at a different queue so the thieves spread out)
// and lane 0 claims from the first one found.  Every failed claim means the
// queue ran out meanwhile (counters only grow), so the search ends.  Returns
// the claimed tile (in every lane) or -1 when all queues are empty.
__device__ int64_t steal(const int64_t* __restrict__ qseg, unsigned int* counters, int nq) {
    const int lane = threadIdx.x;
    const int home = (int)(blockIdx.x % kXcds);
    for (int pass = 0; pass < 2; ++pass) {
        const int span = pass == 0 ? (nq - home + kXcds - 1) / kXcds : nq;
        const int rot = pass == 0 ? (int)(blockIdx.x / kXcds) : (int)blockIdx.x;
        for (;;) {
            int found = -1;
            for (int base = 0; base < span && found < 0; base += 64) {
                const int j = base + lane;
                const int jj = span > 0 ? (j + rot) % span : 0;
                const int v = pass == 0 ? home + kXcds * jj : jj;
                bool has = false;
                if (j < span && v < nq) {
                    const int64_t taken = (int64_t)__hip_atomic_load(&counters[v], __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT);
                    has = taken < qseg[v + 1] - qseg[v];
                }
                const unsigned long long m = __ballot(has);
                if (m) found = __shfl(v, __ffsll((unsigned long long)m) - 1, 64);
            }
            if (found < 0) break;
            int64_t t = -1;
            if (lane == 0) t = claim(qseg, counters, found);
            t = __shfl(t, 0, 64);
            if (t >= 0) return t;
        }
    }
    return -1;
}

__global__ __launch_bounds__(kBlock) void batched_hist_kernel(const Tile* __restrict__ tiles,
                                                             const int64_t* __restrict__ qseg,
                                                             unsigned int* __restrict__ counters,
                                                             int nq) {
    extern __shared__ unsigned long long lds[];
    __shared__ int64_t cur;
    const int me = (int)blockIdx.x;
    int64_t next = -1;  // thread 0: the claim made while the previous tile ran
    if (threadIdx.x == 0) next = claim(qseg, counters, me);
    for (;;) {
        if (threadIdx.x < 64) {  // wave 0
            int64_t t = __shfl(next, 0, 64);
            if (t < 0) t = steal(qseg, counters, nq);
            if (threadIdx.x == 0) cur = t;
        }
        __syncthreads();
        const int64_t t = cur;
        __syncthreads();  // everyone has read `cur` before thread 0 may overwrite it
        if (t < 0) break;  // uniform: every wave of the workgroup leaves
        if (threadIdx.x == 0) next = claim(qseg, counters, me);  // overlaps this tile
        const Tile tile = tiles[t];
        const int64_t nb = tile.bins;
        unsigned long long* h = tile.dst;
        // the remainder's form is chosen per tile, never per element (a branch
        // per element would serialise the loads of the streaming loop)
        const bool pow2 = tile.mod.mask != 0;
        if (tile.mode == kModeGlobal) {
            auto add = [&](int64_t b, int64_t x) {
                if ((uint64_t)b < (uint64_t)nb) global_add(&h[b], (uint64_t)x * (uint64_t)x);
            };
            if (pow2)
                tile_elements<kBlock, true>(tile.src, tile.len, qseg, [&](int64_t x) { add(c_mod_pow2(x, tile.mod), x); });
            else
                tile_elements<kBlock, true>(tile.src, tile.len, qseg, [&](int64_t x) { add(c_mod(x, tile.mod), x); });
            continue;
        }
        const int64_t lo = tile.lo;
        const int width = tile.width;
        // copies of the range at an odd stride (a bin's copies in different
        // banks); a range that fills LDS gets one copy at stride = width
        // (the planner keeps width <= kWords)
        int stride = width | 1;
        int copies = kWords / stride;
        if (copies < 1) {
            copies = 1;
            stride = width;
        } else if (copies > 16) {
            copies = 16;
        }
        const int words = copies * stride;
        for (int i = threadIdx.x; i < words; i += kBlock) lds[i] = 0;
        __syncthreads();
        unsigned long long* mine = lds + (threadIdx.x % copies) * stride;
        if (pow2) {
            // x % m for m = 2^k <= 2^31 in 32-bit arithmetic: the low bits r,
            // kept when x >= 0 or r == 0 (a negative x with r != 0 has the
            // negative C remainder r - m: not a bin)
            const uint32_t mask = (uint32_t)tile.mod.mask, lo32 = (uint32_t)lo, w32 = (uint32_t)width;
            auto f = [&](int64_t x) {
                const uint32_t r = (uint32_t)x & mask;
                const uint32_t i = r - lo32;
                if ((x >= 0 || r == 0) && i < w32) lds_add(&mine[i], (uint64_t)x * (uint64_t)x);
            };
            if (tile.once)
                tile_elements<kBlock, true>(tile.src, tile.len, qseg, f);
            else
                tile_elements<kBlock, false>(tile.src, tile.len, qseg, f);
        } else {
            auto f = [&](int64_t x) {
                const int64_t b = c_mod(x, tile.mod) - lo;
                if ((uint64_t)b < (uint64_t)width) lds_add(&mine[b], (uint64_t)x * (uint64_t)x);
            };
            if (tile.once)
                tile_elements<kBlock, true>(tile.src, tile.len, qseg, f);
            else
                tile_elements<kBlock, false>(tile.src, tile.len, qseg, f);
        }
        __syncthreads();
        if (tile.mode == kModeMerge) {
            for (int b = threadIdx.x; b < width; b += kBlock) {
                unsigned long long s = 0;
                for (int c = 0; c < copies; ++c) s += lds[c * stride + b];
                if (s) global_add(&h[lo + b], s);
            }
        } else {
            // non-temporal stores: the bins are not read again here, and plain
            // stores would take L2 lines from the elements the other range
            // tiles of an instance still read (measured on MI355X: the
            // 4096-instance batch 1.17 -> 1.10-1.12 ms)
            for (int b = threadIdx.x; b < width; b += kBlock) {
                unsigned long long s = 0;
                for (int c = 0; c < copies; ++c) s += lds[c * stride + b];
                __builtin_nontemporal_store(s, &h[lo + b]);
            }
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

struct Group {  // tiles that should run side by side on one XCD (an instance's ranges)
    int64_t cost;
    int64_t first, count;  // in the flat tile list
};

std::once_flag g_attr_once;
hipError_t g_attr_err = hipSuccess;

int64_t tile_cost(const Tile& t) { return t.len + (t.mode != kModeGlobal ? t.width : 0) + kTileOverhead; }

}  // namespace

// Instance k: src[k][0..len[k]) -> dst[k][0..bins[k]).  All arrays are host
// arrays of device pointers / sizes.
int batched_histogram_ptrs(const int64_t* const* src, const int64_t* len, const int64_t* bins,
                           int64_t* const* dst, int64_t count, int policy, hipStream_t s,
                           int64_t* out_tiles) {
    std::vector<Tile> flat;
    std::vector<Group> groups;
    flat.reserve((size_t)count + 64);
    groups.reserve((size_t)count + 64);
    int64_t total_len = 0;
    for (int64_t k = 0; k < count; ++k) total_len += std::max<int64_t>(len[k], 0);
    const int64_t split_at =
        std::max<int64_t>(kMinPiece, total_len / ((int64_t)device_info().num_cus * kPiecesPerQueue));
    std::vector<int64_t> zero;  // (dst pointer, bins) pairs
    TuneSnapshot tune;
    if (policy != 1 && policy != 2) tuned_batch_snapshot(&tune);
    // instances repeat bin counts: keep each count's divisor constants
    int64_t mod_bins[4] = {0, 0, 0, 0};
    ModConst mod_cache[4];
    int mod_next = 0;
    for (int64_t k = 0; k < count; ++k) {
        if (len[k] < 0 || bins[k] < 1 || bins[k] > ((int64_t)1 << 31) || !dst[k] || (len[k] > 0 && !src[k]))
            return set_error(HYD_E_INVALID, "batched_histogram: bad descriptor %lld", (long long)k);
        const int64_t n = len[k], nb = bins[k];
        int slot = -1;
        for (int i = 0; i < 4; ++i)
            if (mod_bins[i] == nb) slot = i;
        if (slot < 0) {
            slot = mod_next;
            mod_next = (mod_next + 1) & 3;
            mod_bins[slot] = nb;
            mod_cache[slot] = make_mod_const(nb);
        }
        const ModConst mod = mod_cache[slot];
        unsigned long long* d = reinterpret_cast<unsigned long long*>(dst[k]);
        int mode;
        if (policy == 1)
            mode = 0;
        else if (policy == 2)
            mode = 1;
        else {
            mode = tune.mode[tune_class_bins(nb)][tune_class_n(n)];
            if (mode < 0) mode = rule_mode(n, nb);
        }
        if (mode == 1) {
            zero.push_back((int64_t)reinterpret_cast<uintptr_t>(dst[k]));
            zero.push_back(nb);
            for (int64_t st = 0; st < n; st += kChunk) {
                flat.push_back(Tile{src[k] + st, d, std::min(kChunk, n - st), nb, mod, 0, 0, kModeGlobal, 1});
                groups.push_back(Group{tile_cost(flat.back()), (int64_t)flat.size() - 1, 1});
            }
            continue;
        }
        // An instance much longer than a queue's share is cut into pieces,
        // each its own LDS tiles whose bins are added into the zeroed
        // destination: one workgroup streams at a fraction of HBM's rate,
        // so one long instance on one workgroup would set the batch's end.
        // An instance with more bins than one range is split only when it is
        // far longer than a queue's share (16x), into at most as many pieces as
        // fill the CUs with its range tiles: each piece streams the elements
        // once per range again and adds every bin of every range with an
        // atomic (measured: slower for instances of up to 2^20 elements;
        // unsplit, one 2^27-element instance with 65536 bins kept 4
        // workgroups busy for 15 ms).
        int64_t pieces = 1;
        if (nb <= kWords) {
            if (n > split_at) pieces = (n + split_at - 1) / split_at;
        } else if (n > 16 * split_at) {
            const int64_t ranges = (nb + kWords - 1) / kWords;
            pieces = std::min<int64_t>((n + split_at - 1) / split_at,
                                       std::max<int64_t>(1, (int64_t)device_info().num_cus / ranges));
        }
        const int32_t once = nb <= kWords ? 1 : 0;  // one range tile reads the elements
        if (pieces > 1) {
            zero.push_back((int64_t)reinterpret_cast<uintptr_t>(dst[k]));
            zero.push_back(nb);
        }
        for (int64_t pc = 0; pc < pieces; ++pc) {
            const int64_t st = n * pc / pieces, en = n * (pc + 1) / pieces;
            const int64_t pfirst = (int64_t)flat.size();
            int64_t cost = 0;
            for (int64_t lo = 0; lo < nb; lo += kWords) {
                const int32_t w = (int32_t)std::min<int64_t>(kWords, nb - lo);
                flat.push_back(Tile{src[k] + st, d, en - st, nb, mod, (int32_t)lo, w,
                                    pieces > 1 ? kModeMerge : kModeOwner, once});
                cost += tile_cost(flat.back());
            }
            groups.push_back(Group{cost, pfirst, (int64_t)flat.size() - pfirst});
        }
    }
    const int nq = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)flat.size(), device_info().num_cus));
    // Gang scheduling: teams of kTeam workgroups on one XCD (queue q runs on
    // XCD q % 8) get aligned queues - job j hands each member one tile: the
    // range tiles of one instance (which stream the same elements and so share
    // them through L2 only while they run side by side), or a bundle of
    // single tiles of similar length.  Jobs go longest first onto the
    // least-loaded team (LPT).  Batches too small for full teams use teams of
    // one workgroup.
    std::vector<std::vector<int64_t>> qtiles((size_t)nq);
    {
        const int per_x = nq / kXcds;
        const int team = (nq % kXcds == 0 && per_x >= kTeam && per_x % kTeam == 0) ? kTeam : 1;
        std::vector<std::vector<int>> teams;  // member queues
        for (int x = 0; x < std::min(kXcds, nq); ++x)
            for (int q0 = x; q0 < nq; q0 += kXcds * team) {
                std::vector<int> m;
                for (int i = 0; i < team && q0 + kXcds * i < nq; ++i) m.push_back(q0 + kXcds * i);
                teams.push_back(m);
            }
        // Jobs hold at most `team` tile indices inline (no allocation per
        // job), and the least-loaded team comes from a heap ordered by
        // (load, team index): the same choice as scanning the teams for the
        // first minimal load, in O(log teams) - the planner runs on the host
        // before the launch, so its time adds to the batch's (4096
        // instances: 0.45 ms with a scan and per-job vectors).
        struct Job {
            int64_t cost;
            int32_t n;
            int64_t tiles[kTeam];
        };
        std::vector<int64_t> tcost(flat.size());
        for (size_t t = 0; t < flat.size(); ++t) tcost[t] = tile_cost(flat[t]);
        std::vector<Job> jobs;
        jobs.reserve(flat.size() / (size_t)team + groups.size() + 1);
        std::vector<int64_t> singles;
        singles.reserve(groups.size());
        for (const Group& g : groups) {
            if (g.count == 1) {
                singles.push_back(g.first);
                continue;
            }
            for (int64_t i = 0; i < g.count; i += team) {
                Job jb{0, 0, {}};
                for (int64_t k = i; k < std::min<int64_t>(g.count, i + team); ++k) {
                    jb.tiles[jb.n++] = g.first + k;
                    jb.cost = std::max(jb.cost, tcost[(size_t)(g.first + k)]);
                }
                jobs.push_back(jb);
            }
        }
        std::stable_sort(singles.begin(), singles.end(),
                         [&](int64_t a, int64_t b) { return tcost[(size_t)a] > tcost[(size_t)b]; });
        for (size_t i = 0; i < singles.size(); i += (size_t)team) {
            Job jb{0, 0, {}};
            for (size_t k = i; k < std::min(singles.size(), i + (size_t)team); ++k) {
                jb.tiles[jb.n++] = singles[k];
                jb.cost = std::max(jb.cost, tcost[(size_t)singles[k]]);
            }
            jobs.push_back(jb);
        }
        std::stable_sort(jobs.begin(), jobs.end(), [](const Job& a, const Job& b) { return a.cost > b.cost; });
        for (auto& q : qtiles) q.reserve(flat.size() / (size_t)nq + 8);
        std::priority_queue<std::pair<int64_t, int>, std::vector<std::pair<int64_t, int>>,
                            std::greater<std::pair<int64_t, int>>>
            heap;
        for (int t = 0; t < (int)teams.size(); ++t) heap.push({0, t});
        for (const Job& jb : jobs) {
            const std::pair<int64_t, int> top = heap.top();
            heap.pop();
            const std::vector<int>& members = teams[(size_t)top.second];
            for (int k = 0; k < jb.n && k < (int)members.size(); ++k) qtiles[(size_t)members[(size_t)k]].push_back(jb.tiles[k]);
            heap.push({top.first + jb.cost, top.second});
        }
    }
    std::vector<int64_t> qseg((size_t)nq + 1);
    for (int q = 0; q < nq; ++q) qseg[q + 1] = qseg[q] + (int64_t)qtiles[q].size();
    if (qseg[(size_t)nq] != (int64_t)flat.size())
        return set_error(HYD_E_INVALID, "batched_histogram: planner placed %lld of %zu tiles",
                         (long long)qseg[(size_t)nq], flat.size());
    if (out_tiles) *out_tiles = (int64_t)flat.size();

    // descriptor block: counters (zeroed), queue boundaries, tiles, zero list
    const size_t cnt_bytes = ((size_t)nq * 4 + 255) & ~(size_t)255;
    const size_t seg_bytes = (((size_t)nq + 1) * 8 + 255) & ~(size_t)255;
    const size_t tile_bytes = flat.size() * sizeof(Tile);
    const size_t zero_bytes = zero.size() * sizeof(int64_t);
    const size_t bytes = cnt_bytes + seg_bytes + tile_bytes + zero_bytes;
    unsigned char* ws = static_cast<unsigned char*>(stream_workspace(s, bytes));
    if (!ws) return set_error(HYD_E_OOM, "batched_histogram: workspace allocation failed");
    HYD_TRY(stage_h2d(s, ws, bytes, [&](unsigned char* q) {
        memset(q, 0, cnt_bytes);
        memcpy(q + cnt_bytes, qseg.data(), qseg.size() * 8);
        Tile* t = reinterpret_cast<Tile*>(q + cnt_bytes + seg_bytes);
        for (int qq = 0; qq < nq; ++qq)
            for (int64_t ti : qtiles[qq]) *t++ = flat[(size_t)ti];
        memcpy(q + cnt_bytes + seg_bytes + tile_bytes, zero.data(), zero_bytes);
    }));
    unsigned int* counters = reinterpret_cast<unsigned int*>(ws);
    const int64_t* d_seg = reinterpret_cast<const int64_t*>(ws + cnt_bytes);
    const Tile* d_tiles = reinterpret_cast<const Tile*>(ws + cnt_bytes + seg_bytes);
    const int64_t* d_zero = reinterpret_cast<const int64_t*>(ws + cnt_bytes + seg_bytes + tile_bytes);
    if (!zero.empty()) {
        const int64_t nr = (int64_t)zero.size() / 2;
        zero_ranges_kernel<<<(unsigned)std::min<int64_t>(nr, 4096), 256, 0, s>>>(d_zero, nr);
        HYD_CHECK_LAUNCH();
    }
    if (flat.empty()) return HYD_OK;
    const int lds = kWords * 8;
    std::call_once(g_attr_once, [&] {
        g_attr_err = hipFuncSetAttribute((const void*)batched_hist_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    });
    HYD_TRY_HIP(g_attr_err);
    batched_hist_kernel<<<(unsigned)nq, kBlock, lds, s>>>(d_tiles, d_seg, counters, nq);
    HYD_CHECK_LAUNCH();
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
