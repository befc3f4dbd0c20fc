// csv.hip - CSV text <-> int64 / binary64 arrays on the GPU: the read_array /
// read_array_float / write_array / write_array_float runtime of the
// reference (runtime/hydrangea_runtime.h:549-594, runtime/hyd_write_csv.c),
// with the number conversions of csv_num.h.
//
// Parse (text in HBM -> values):
//   1. count: each workgroup stages 8 KiB of text (plus halos) in LDS; a word
//      is a maximal run of non-gap characters (gap = ',' or isspace).  The
//      thread whose 32-byte slice holds a word's first character checks the
//      gap before it and scans the word's numbers (a word holds several when
//      they abut: "1-2" is 1 then -2).  Per tile: the number count and the
//      first failure (local index).
//   2. an exclusive scan of the tile counts gives each tile's first element.
//   3. values: the same scan again, converting each number and storing
//      element g < count.  A decimal the 128-bit table cannot round with
//      certainty is marked in a bitmap, with its text position in its slot.
//   4. exact: big-integer rounding of the marked elements (rare).
//   5. status: the first failure over all tiles, or, when fewer numbers than
//      elements exist, what the reference's loop reports at the end.
//
// Format (values -> text in HBM): per element the digits (fp64: 17
// significant, table path, exact path for undecided roundings), per
// 1024-element block the text length, a scan of the block lengths, then each
// block writes its text into LDS and from there to HBM with 16-byte stores.
//
// The gap rule between two numbers follows the reference's loop: after an
// element it skips ',', ' ', '\n', '\r' (fgetc), then fscanf skips isspace,
// so a gap is [, \n\r]* followed by [ \t\n\v\f\r]*: a ',' after a '\t', '\v'
// or '\f' fails the next element, and before the first element only
// isspace is skipped.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <errno.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hyd_hip_runtime.h"
#include "csv_num.h"
#include "hyd_internal.hpp"

namespace hyd {
int scan_i64(const int64_t* in, int64_t n, int64_t init, int64_t* out, bool inclusive, bool reverse,
             hipStream_t s, int64_t* tile_ws);
int64_t scan_workspace_words(int64_t n);
hipStream_t export_stream();
int host_sync_enabled();
void mark_device_written(hyd_array_t* arr);

namespace {

using namespace csv;

constexpr int kPThreads = 256;
constexpr int kWaves = kPThreads / 64;
constexpr int kSlice = 32;                          // text bytes per thread
constexpr int kTile = kPThreads * kSlice;           // 8 KiB per workgroup
constexpr int kHaloLo = 64, kHaloHi = 448;          // LDS beyond the tile
constexpr int kBuf = kHaloLo + kTile + kHaloHi;     // 8704 bytes, a multiple of 16
constexpr int kChunks = kBuf / 16;

// ---- block helpers --------------------------------------------------------
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* total, int64_t* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    int64_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int64_t t = sh[w];
        if (w < wid) off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__device__ __forceinline__ int64_t block_min(int64_t v, int64_t* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const int64_t y = __shfl_xor(v, d, 64);
        v = y < v ? y : v;
    }
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int64_t m = sh[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) m = sh[w] < m ? sh[w] : m;
    __syncthreads();
    return m;
}

// Sum of v over the workgroup, valid in thread 0 (one barrier; |sum| < 2^31).
__device__ __forceinline__ int block_sum_i32(int v, int* sh) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += sh[w];
    return t;
}

// Stages text[base, base + kBuf) into LDS (bytes outside [0, n) are left
// unused: the accessor checks the range first).
template <bool ALIGNED>
__device__ __forceinline__ void stage_tile(const unsigned char* __restrict__ text, int64_t n, int64_t base,
                                           unsigned char* buf) {
    for (int j = threadIdx.x; j < kChunks; j += kPThreads) {
        const int64_t p = base + 16 * (int64_t)j;
        if (ALIGNED && p >= 0 && p + 16 <= n) {
            *reinterpret_cast<uint4*>(buf + 16 * j) = *reinterpret_cast<const uint4*>(text + p);
        } else {
            for (int b = 0; b < 16; ++b)
                if (p + b >= 0 && p + b < n) buf[16 * j + b] = text[p + b];
        }
    }
    __syncthreads();
}

// What the LDS-only accessor returns for a text byte that is not staged: no
// character class matches it, so a scan stops there, and the caller redoes
// the word with the full accessor.
constexpr int kSpill = 0x100;

// Backward check of the gap that ends at p - 1 (p: a word's first char):
// 1 = a gap the reference's loop skips, 0 = not, 2 = ran out of staged text.
template <class At>
__device__ __forceinline__ int gap_state(const At& at, int64_t p) {
    bool comma_seen = false;  // a ',' lies between here and the word
    for (int64_t i = p - 1;; --i) {
        const int c = at(i);
        if (c == kSpill) return 2;
        if (c < 0) return comma_seen ? 0 : 1;  // start of text: only isspace allowed
        if (!is_gap(c)) return 1;              // end of the previous word
        if (c == ',') comma_seen = true;
        else if (comma_seen && !is_sep(c)) return 0;  // '\t' '\v' '\f' before a ','
    }
}

struct NullSink {
    int64_t qadj = 0;
    __device__ __forceinline__ void digit(int, bool, bool) {}
};

template <bool FLOAT, class At>
__device__ __forceinline__ int64_t token_len(const At& at, int64_t p) {
    if (FLOAT) {
        NullSink s;
        return scan_f64(at, p, s).len;
    }
    return scan_i64(at, p, (int64_t*)nullptr);
}

constexpr int kMaxWords = kTile / 2 + 1;  // a word start follows a gap (or is the text's first byte)

// Word starts among 32 bytes (qa, qb: bytes 0-15, 16-31): bit k set when byte
// k is not a gap and byte k - 1 is one (before_gap: the byte before byte 0 is
// a gap or the text's start).  Bytes from `valid` on are past the text.
__device__ __forceinline__ uint32_t word_starts32(uint4 qa, uint4 qb, bool before_gap, int valid) {
    const uint32_t gap = gap_mask8(((uint64_t)qa.y << 32) | qa.x) | (gap_mask8(((uint64_t)qa.w << 32) | qa.z) << 8) |
                         (gap_mask8(((uint64_t)qb.y << 32) | qb.x) << 16) |
                         (gap_mask8(((uint64_t)qb.w << 32) | qb.z) << 24);
    const uint32_t in = valid >= 32 ? 0xFFFFFFFFu : ((1u << valid) - 1);
    return ~gap & ((gap << 1) | (before_gap ? 1u : 0u)) & in;
}

// Stages the tile, then lists the words that start in it: starts[0, words)
// (tile offsets, in text order).  A byte-parallel pass: each thread marks the
// starts among its kSlice bytes, a block scan places them.
#define HYD_TILE_WORDS                                                                  \
    __shared__ __attribute__((aligned(16))) unsigned char buf[kBuf];                    \
    __shared__ uint16_t starts[kMaxWords];                                              \
    __shared__ int64_t sh[kWaves];                                                      \
    const int64_t t0 = (int64_t)blockIdx.x * kTile;                                     \
    const int64_t base = t0 - kHaloLo;                                                  \
    const int64_t lo = base < 0 ? 0 : base;            /* staged: text[lo, lim) */      \
    const int64_t lim = n < base + kBuf ? n : base + kBuf;                              \
    const int64_t stage_end = lim < n ? lim : INT64_MAX; /* unstaged text from here */  \
    stage_tile<ALIGNED>(text, n, base, buf);                                            \
    auto at = [&](int64_t p) -> int {  /* any position */                              \
        if ((uint64_t)(p - lo) < (uint64_t)(lim - lo)) return buf[p - base];            \
        return (p < 0 || p >= n) ? -1 : text[p];                                        \
    };                                                                                  \
    auto atl = [&](int64_t p) -> int {  /* staged bytes only, without a branch */      \
        const bool in = (uint64_t)(p - lo) < (uint64_t)(lim - lo);                      \
        const int c = buf[in ? p - base : 0];                                           \
        return in ? c : ((p < 0 || p >= n) ? -1 : kSpill);                              \
    };                                                                                  \
    int words;                                                                          \
    {                                                                                   \
        const int64_t s0 = t0 + (int64_t)threadIdx.x * kSlice;                          \
        const int64_t rest = n - s0;                                                    \
        const int valid = rest <= 0 ? 0 : (rest < kSlice ? (int)rest : kSlice);         \
        const uint4 qa = *reinterpret_cast<const uint4*>(buf + (s0 - base));            \
        const uint4 qb = *reinterpret_cast<const uint4*>(buf + (s0 - base) + 16);       \
        const int prev = at(s0 - 1);                                                    \
        uint32_t mask = word_starts32(qa, qb, prev < 0 || is_gap(prev), valid);         \
        int64_t total;                                                                  \
        int off = (int)block_excl_scan(__builtin_popcount(mask), &total, sh);           \
        while (mask) {                                                                  \
            const int k = __builtin_ctz(mask);                                          \
            mask &= mask - 1;                                                           \
            starts[off++] = (uint16_t)(threadIdx.x * kSlice + k);                       \
        }                                                                               \
        words = (int)total;                                                             \
        __syncthreads();                                                                \
    }

// Scans the word at p: how many numbers it holds (those before a failure),
// and whether one failed (or the gap before the word is not one the
// reference's loop skips).  first(q) scans the first number and returns its
// length; the others are scanned for their lengths only.  *redo is set when
// the scan may have met an unstaged byte (the LDS-only accessor): a number
// scan reads at most 8 bytes past where it stops.
template <bool FLOAT, class At, class First>
__device__ __forceinline__ int64_t scan_word(const At& at, int64_t p, int64_t stage_end, bool* fail, bool* redo,
                                             const First& first) {
    *fail = false;
    *redo = false;
    const int g = gap_state(at, p);
    if (g != 1) {
        *redo = g == 2;
        *fail = g == 0;
        return 0;
    }
    int64_t cnt = 0, q = p;
    for (;;) {
        const int64_t len = cnt == 0 ? first(q) : token_len<FLOAT>(at, q);
        if (len < 0) {
            *fail = true;
            *redo = q + 8 >= stage_end;
            return cnt;
        }
        ++cnt;
        q += len;
        const int c = at(q);
        if (c < 0 || is_gap(c) || c == kSpill) {
            *redo = q + 8 >= stage_end;
            return cnt;
        }
    }
}

// Words that start in each tile (no number scan): the element index of
// every word when each holds one number, the common case.
template <bool ALIGNED>
__global__ __launch_bounds__(kPThreads) void csv_words_kernel(const unsigned char* __restrict__ text, int64_t n,
                                                              int64_t* __restrict__ tile_cnt) {
    __shared__ int sh[kWaves];
    const int64_t s0 = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kSlice;
    uint32_t wv[8];
    if (ALIGNED && s0 + kSlice <= n) {
        const uint4 qa = *reinterpret_cast<const uint4*>(text + s0);
        const uint4 qb = *reinterpret_cast<const uint4*>(text + s0 + 16);
        wv[0] = qa.x; wv[1] = qa.y; wv[2] = qa.z; wv[3] = qa.w;
        wv[4] = qb.x; wv[5] = qb.y; wv[6] = qb.z; wv[7] = qb.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t v = 0;
            for (int b = 0; b < 4; ++b)
                if (s0 + 4 * k + b < n) v |= (uint32_t)text[s0 + 4 * k + b] << (8 * b);
            wv[k] = v;
        }
    }
    const int64_t rest = n - s0;
    const int valid = rest <= 0 ? 0 : (rest < kSlice ? (int)rest : kSlice);
    const int prev = (s0 > 0 && s0 - 1 < n) ? text[s0 - 1] : -1;
    const uint4 qa = make_uint4(wv[0], wv[1], wv[2], wv[3]), qb = make_uint4(wv[4], wv[5], wv[6], wv[7]);
    const int cnt = __builtin_popcount(word_starts32(qa, qb, prev < 0 || is_gap(prev), valid));
    const int total = block_sum_i32(cnt, sh);  // at most kTile
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = total;
}

template <bool FLOAT, bool ALIGNED>
__global__ __launch_bounds__(kPThreads) void csv_count_kernel(const unsigned char* __restrict__ text, int64_t n,
                                                              int64_t* __restrict__ tile_cnt,
                                                              int64_t* __restrict__ tile_err) {
    HYD_TILE_WORDS
    int64_t carry = 0, err = INT64_MAX;
    for (int r0 = 0; r0 < words; r0 += kPThreads) {  // one word per thread per round
        const int w = r0 + (int)threadIdx.x;
        int64_t cnt = 0;
        bool fail = false;
        if (w < words) {
            const int64_t p = t0 + starts[w];
            bool redo;
            cnt = scan_word<FLOAT>(atl, p, stage_end, &fail, &redo,
                                   [&](int64_t q) { return token_len<FLOAT>(atl, q); });
            if (redo)  // a word longer than the staged halo, or a gap reaching before it
                cnt = scan_word<FLOAT>(at, p, INT64_MAX, &fail, &redo,
                                       [&](int64_t q) { return token_len<FLOAT>(at, q); });
        }
        int64_t total;
        const int64_t off = block_excl_scan(cnt, &total, sh);
        if (fail) err = min(err, carry + off + cnt);
        carry += total;
    }
    err = block_min(err, sh);
    if (threadIdx.x == 0) {
        tile_cnt[blockIdx.x] = carry;
        tile_err[blockIdx.x] = err == INT64_MAX ? -1 : err;
    }
}

// The value of the number at p: its bits (int64 value, or binary64 bits)
// and whether the table rounding decided it; returns its length (-1: none).
template <bool FLOAT, class At>
__device__ __forceinline__ int64_t read_one(const At& at, int64_t p, uint64_t* bits, bool* decided) {
    if (FLOAT) {
        const FloatTok t = scan_f64_fast(at, p);
        if (t.len >= 0) *decided = tok_to_bits(t, bits);
        return t.len;
    }
    int64_t v;
    const int64_t len = scan_i64(at, p, &v);
    *bits = (uint64_t)v;
    *decided = true;
    return len;
}

// Stores a number read by read_one as element g (an undecided binary64 as
// its text position, marked for the exact kernel).
__device__ __forceinline__ void store_one(int64_t g, int64_t p, uint64_t bits, bool decided, int64_t count,
                                          int64_t* __restrict__ out, uint32_t* __restrict__ exact_bits,
                                          unsigned long long* __restrict__ exact_cnt) {
    if (g >= count) return;
    if (decided) {
        out[g] = (int64_t)bits;
    } else {
        out[g] = p;  // the exact kernel rereads the text there
        atomicOr(&exact_bits[g >> 5], 1u << (g & 31));
        atomicAdd(exact_cnt, 1ULL);
    }
}

// Values.  ONE_PER_WORD: tile_base counts words (csv_words_kernel), which
// equals the numbers' indices when every word holds exactly one number after
// a gap the reference skips; a word where that fails sets *irregular and the
// caller redoes the text with counted numbers (ONE_PER_WORD = false, tile_base
// from csv_count_kernel).
template <bool FLOAT, bool ALIGNED, bool ONE_PER_WORD>
__global__ __launch_bounds__(kPThreads) void csv_value_kernel(const unsigned char* __restrict__ text, int64_t n,
                                                              const int64_t* __restrict__ tile_base,
                                                              const int64_t* __restrict__ tile_err,
                                                              int64_t count, int64_t* __restrict__ out,
                                                              uint32_t* __restrict__ exact_bits,
                                                              unsigned long long* __restrict__ exact_cnt,
                                                              unsigned long long* __restrict__ err_enc,
                                                              unsigned long long* __restrict__ irregular) {
    const int64_t g0 = tile_base[blockIdx.x];
    if (g0 >= count) return;  // block-uniform: no element of this tile is kept
    if (!ONE_PER_WORD && threadIdx.x == 0 && tile_err[blockIdx.x] >= 0)
        atomicMax(err_enc, (unsigned long long)(INT64_MAX - (g0 + tile_err[blockIdx.x])));
    HYD_TILE_WORDS
    // 8 staged bytes from q (q - base + 16 <= kBuf)
    auto rd8 = [&](int64_t q) -> uint64_t {
        const int o = (int)(q - base);
        const int a = o & ~7, sh = (o & 7) * 8;
        const uint64_t lo8 = *reinterpret_cast<const uint64_t*>(buf + a);
        const uint64_t hi8 = *reinterpret_cast<const uint64_t*>(buf + a + 8);
        return sh ? (lo8 >> sh) | (hi8 << (64 - sh)) : lo8;
    };
    // the word's first number, read in the counting scan itself: the 8-byte
    // path when its bytes are staged, else the character scanner
    auto first_lds = [&](int64_t q, uint64_t* bits, bool* dec) -> int64_t {
        if (FLOAT && q + kSwarSpan + 8 <= lim) {
            FloatTok t;
            if (scan_dec_swar(rd8, atl, q, &t)) {
                *dec = tok_to_bits(t, bits);
                return t.len;
            }
        }
        return read_one<FLOAT>(atl, q, bits, dec);
    };
    int64_t carry = g0;
    for (int r0 = 0; r0 < words; r0 += kPThreads) {
        const int w = r0 + (int)threadIdx.x;
        int64_t cnt = 0, p = 0, len1 = 0;
        uint64_t bits1 = 0;
        bool dec1 = true, fail = false;
        const bool mine = w < words && (!ONE_PER_WORD || g0 + w < count);
        if (mine) {
            p = t0 + starts[w];
            bool redo;
            cnt = scan_word<FLOAT>(atl, p, stage_end, &fail, &redo,
                                   [&](int64_t q) { return len1 = first_lds(q, &bits1, &dec1); });
            if (redo)
                cnt = scan_word<FLOAT>(at, p, INT64_MAX, &fail, &redo,
                                       [&](int64_t q) { return len1 = read_one<FLOAT>(at, q, &bits1, &dec1); });
        }
        if (ONE_PER_WORD) {  // word w is element g0 + w
            if (mine) {
                if (cnt != 1 || fail) atomicOr(irregular, 1ULL);
                else store_one(g0 + w, p, bits1, dec1, count, out, exact_bits, exact_cnt);
            }
            continue;
        }
        int64_t total;
        const int64_t g = carry + block_excl_scan(cnt, &total, sh);
        if (cnt > 0) {
            store_one(g, p, bits1, dec1, count, out, exact_bits, exact_cnt);
            int64_t q = p + len1;
            for (int64_t j = 1; j < cnt; ++j) {  // numbers that abut ("1-2"): rare
                uint64_t b;
                bool d;
                const int64_t l = read_one<FLOAT>(at, q, &b, &d);
                store_one(g + j, q, b, d, count, out, exact_bits, exact_cnt);
                q += l;
            }
        }
        carry += total;
    }
}

// Exact rounding of the marked decimals (few; runs on a small grid).
__global__ __launch_bounds__(64) void csv_exact_kernel(const unsigned char* __restrict__ text, int64_t n,
                                                       int64_t count, const uint32_t* __restrict__ exact_bits,
                                                       const unsigned long long* __restrict__ exact_cnt,
                                                       int64_t* __restrict__ out) {
    if (*exact_cnt == 0) return;
    auto at = [&](int64_t p) -> int { return (p < 0 || p >= n) ? -1 : text[p]; };
    const int64_t words = (count + 31) >> 5;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint32_t b = exact_bits[w];
        while (b) {
            const int bit = __builtin_ctz(b);
            b &= b - 1;
            const int64_t g = w * 32 + bit;
            const int64_t p = out[g];
            const FloatTok t = scan_f64_fast(at, p);
            uint64_t cand;
            tok_to_bits(t, &cand);
            Big D, L, R;
            out[g] = (int64_t)dec_exact_bits(at, p, cand, D, L, R);
        }
    }
}

// status[0] = first failing element (-1: none below count), status[1] = kind,
// status[2] = numbers found (capped by the first failure), status[3] = exact.
// err_enc = INT64_MAX - (first failing element over all tiles), 0: none.
__global__ __launch_bounds__(kPThreads) void csv_status_kernel(const unsigned char* __restrict__ text, int64_t n,
                                                               int64_t count, int64_t tiles,
                                                               const int64_t* __restrict__ tile_cnt,
                                                               const int64_t* __restrict__ tile_base,
                                                               const unsigned long long* __restrict__ exact_cnt,
                                                               const unsigned long long* __restrict__ err_enc,
                                                               int64_t* __restrict__ status) {
    __shared__ int64_t sh[kWaves];
    int64_t e = *err_enc ? INT64_MAX - (int64_t)*err_enc : INT64_MAX;
    int64_t found = tiles ? tile_base[tiles - 1] + tile_cnt[tiles - 1] : 0;
    if (e < found) found = e;  // numbers after a failure are not the reference's elements
    int64_t kind = 0;
    if (e != INT64_MAX) {
        kind = HYDHIP_CSV_PARSE_ERROR;
    } else if (found < count) {
        e = found;
        if (found == 0) {
            kind = HYDHIP_CSV_PARSE_ERROR;  // the first fscanf meets the end
        } else {
            // the text after the last number is all gap (no word failed):
            // the separator loop reaches the end unless a '\t' '\v' '\f'
            // stops it, after which fscanf meets the end or a ','.  Scan
            // back from the end to the last non-gap character.
            int64_t bad = 0;
            for (int64_t c0 = n - 1; c0 >= 0; c0 -= kPThreads) {
                const int64_t p = c0 - threadIdx.x;
                const int c = p >= 0 ? text[p] : ',';
                const bool word = p >= 0 && !is_gap(c);
                const bool tab = p >= 0 && is_gap(c) && !is_sep(c);
                const int64_t last = -block_min(word ? -p : 1, sh);  // -1: none in this chunk
                bad |= -block_min((tab && p > last) ? -1 : 0, sh);
                if (last >= 0) break;
            }
            kind = bad ? HYDHIP_CSV_PARSE_ERROR : HYDHIP_CSV_TOO_FEW;
        }
    }
    if (e >= count) {
        e = -1;
        kind = 0;
    }
    if (threadIdx.x == 0) {
        status[0] = e;
        status[1] = kind;
        status[2] = found;
        status[3] = (int64_t)*exact_cnt;
    }
}

// ---- format ---------------------------------------------------------------
constexpr int kFThreads = 256;
constexpr int kFPer = 4;
constexpr int kFBlock = kFThreads * kFPer;  // elements per workgroup
constexpr int kMaxText = 25;               // "-1.2345678901234567e-308" + separator
constexpr int kStage = kFBlock * kMaxText + 32;

struct Dig {
    uint64_t N;  // 17 digits; the bits for zero, inf, nan
    int32_t X;
    int32_t flags;  // 1 = special (zero, inf, nan), 2 = negative
};

// The %.17g digits of value b by the table path; *decided = false when the
// exact kernel must settle them (N, X then hold its starting point).  FORCE
// (a test mode) leaves every finite nonzero value to the exact kernel.
template <bool FORCE = false>
__device__ __forceinline__ Dig table_digits(uint64_t b, bool* decided) {
    Dig d;
    *decided = true;
    if (special_len(b)) {
        d.N = b;
        d.X = 0;
        d.flags = 1;
        return d;
    }
    uint64_t N;
    int X;
    *decided = f64_to_dec17<FORCE>(b & ~kSignBit, &N, &X);
    d.N = N;
    d.X = X;
    d.flags = (b >> 63) ? 2 : 0;
    return d;
}

__device__ __forceinline__ int dig_len(const Dig& d) {
    return (d.flags & 1) ? special_len(d.N) : g17_len(g17_make(d.flags & 2, d.N, d.X));
}

// A finite nonzero value's table digits in one word for the write kernel:
// N - 10^16 (< 9 * 10^16 < 2^57) in the low 57 bits, X - dec17_x0(bits)
// (0, 1 or 2) above them.  Zero, inf and nan are told by the value's bits.
__device__ __forceinline__ uint64_t pack_digits(uint64_t b, const Dig& d) {
    return (d.flags & 1) ? 0 : (d.N - kTen16) | ((uint64_t)(d.X - dec17_x0(b & ~kSignBit)) << 57);
}
__device__ __forceinline__ Dig unpack_digits(uint64_t b, uint64_t packed) {
    Dig d;
    if (special_len(b)) {
        d.N = b;
        d.X = 0;
        d.flags = 1;
    } else {
        d.N = (packed & ((1ULL << 57) - 1)) + kTen16;
        d.X = dec17_x0(b & ~kSignBit) + (int)(packed >> 57);
        d.flags = (b >> 63) ? 2 : 0;
    }
    return d;
}

// A thread's kFPer consecutive words a[i0, i0 + kFPer) (i0 a multiple of
// kFPer): two 16-byte accesses when all lie in range and the array is 16-byte
// aligned, so that a wave's loads and stores cover whole lines.
__device__ __forceinline__ void load_group(const uint64_t* a, int64_t i0, int64_t n, uint64_t* v) {
    if (i0 + kFPer <= n && (reinterpret_cast<uintptr_t>(a) & 15) == 0) {
        const ulonglong2 p0 = *reinterpret_cast<const ulonglong2*>(a + i0);
        const ulonglong2 p1 = *reinterpret_cast<const ulonglong2*>(a + i0 + 2);
        v[0] = p0.x;
        v[1] = p0.y;
        v[2] = p1.x;
        v[3] = p1.y;
    } else {
#pragma unroll
        for (int k = 0; k < kFPer; ++k) v[k] = i0 + k < n ? a[i0 + k] : 0;
    }
}
__device__ __forceinline__ void store_group(uint64_t* a, int64_t i0, int64_t n, const uint64_t* v) {
    if (i0 + kFPer <= n && (reinterpret_cast<uintptr_t>(a) & 15) == 0) {
        *reinterpret_cast<ulonglong2*>(a + i0) = make_ulonglong2(v[0], v[1]);
        *reinterpret_cast<ulonglong2*>(a + i0 + 2) = make_ulonglong2(v[2], v[3]);
    } else {
#pragma unroll
        for (int k = 0; k < kFPer; ++k)
            if (i0 + k < n) a[i0 + k] = v[k];
    }
}
static_assert(kFPer == 4, "load_group moves 4 words");

// Text length of every 1024-value block (with separators).  fp64 values the
// table path cannot round with certainty are marked (their provisional length
// is corrected by fmt_exact_kernel).  The fp64 digits go to `packed` (one
// word per value) so that the write kernel does not compute them again.
template <bool FLOAT, bool FORCE>
__global__ __launch_bounds__(kFThreads) void fmt_len_kernel(const uint64_t* __restrict__ vals, int64_t n,
                                                            int64_t* __restrict__ block_len,
                                                            uint32_t* __restrict__ exact_bits,
                                                            unsigned long long* __restrict__ exact_cnt,
                                                            uint64_t* __restrict__ packed) {
    static_assert(kFThreads == kPThreads, "block_sum_i32 sums kPThreads threads");
    __shared__ int sh[kWaves];
    const int64_t i0 = (int64_t)blockIdx.x * kFBlock + threadIdx.x * kFPer;
    int64_t len = 0;
    uint64_t b[kFPer], pk[kFPer];
    load_group(vals, i0, n, b);
#pragma unroll
    for (int k = 0; k < kFPer; ++k) {
        const int64_t i = i0 + k;
        pk[k] = 0;
        if (i >= n) continue;
        if (FLOAT) {
            bool decided;
            const Dig d = table_digits<FORCE>(b[k], &decided);
            pk[k] = pack_digits(b[k], d);
            len += dig_len(d) + 1;
            if (!decided) {
                atomicOr(&exact_bits[i >> 5], 1u << (i & 31));
                atomicAdd(exact_cnt, 1ULL);
            }
        } else {
            len += i64_len((int64_t)b[k]) + 1;
        }
    }
    if (FLOAT) store_group(packed, i0, n, pk);
    const int total = block_sum_i32((int)len, sh);  // at most kFBlock * kMaxText bytes
    if (threadIdx.x == 0) block_len[blockIdx.x] = total;
}

// Settles the marked values exactly: their digits go to exact_dig[i] (a
// sparse table: only marked entries are written or read) and their block's
// length is corrected.
__global__ __launch_bounds__(64) void fmt_exact_kernel(const uint64_t* __restrict__ vals, int64_t n,
                                                       Dig* __restrict__ exact_dig,
                                                       const uint32_t* __restrict__ exact_bits,
                                                       const unsigned long long* __restrict__ exact_cnt,
                                                       int64_t* __restrict__ block_len) {
    if (*exact_cnt == 0) return;
    const int64_t words = (n + 31) >> 5;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint32_t bm = exact_bits[w];
        while (bm) {
            const int bit = __builtin_ctz(bm);
            bm &= bm - 1;
            const int64_t i = w * 32 + bit;
            bool decided;
            Dig d = table_digits<true>(vals[i], &decided);  // the starting point the length kernel counted
            const int before = dig_len(d);
            Big D, L, R;
            uint64_t N;
            int X;
            dec17_exact(vals[i] & ~kSignBit, d.N, d.X, &N, &X, D, L, R);
            d.N = N;
            d.X = X;
            exact_dig[i] = d;
            atomicAdd(reinterpret_cast<unsigned long long*>(&block_len[i / kFBlock]),
                      (unsigned long long)(int64_t)(dig_len(d) - before));
        }
    }
}

template <bool FLOAT>
__global__ __launch_bounds__(kFThreads) void fmt_write_kernel(const uint64_t* __restrict__ vals, int64_t n,
                                                              int64_t row_len, const int64_t* __restrict__ block_off,
                                                              unsigned char* __restrict__ text,
                                                              const Dig* __restrict__ exact_dig,
                                                              const uint32_t* __restrict__ exact_bits,
                                                              const unsigned long long* __restrict__ exact_cnt,
                                                              const uint64_t* __restrict__ packed) {
    __shared__ int64_t sh[kWaves];
    __shared__ __attribute__((aligned(16))) unsigned char stage[kStage];
    const int64_t i0 = (int64_t)blockIdx.x * kFBlock + threadIdx.x * kFPer;
    const bool any_exact = FLOAT && *exact_cnt != 0;
    Dig dg[kFPer];
    int lens[kFPer];
    int64_t len = 0;
    uint64_t b[kFPer], pk[kFPer];
    load_group(vals, i0, n, b);
    if (FLOAT) load_group(packed, i0, n, pk);
#pragma unroll
    for (int k = 0; k < kFPer; ++k) {
        const int64_t i = i0 + k;
        lens[k] = 0;
        if (i >= n) continue;
        if (FLOAT) {
            dg[k] = unpack_digits(b[k], pk[k]);
            if (any_exact && ((exact_bits[i >> 5] >> (i & 31)) & 1)) dg[k] = exact_dig[i];
            lens[k] = dig_len(dg[k]) + 1;
        } else {
            lens[k] = i64_len((int64_t)b[k]) + 1;
        }
        len += lens[k];
    }
    int64_t total;
    const int64_t off = block_excl_scan(len, &total, sh);
    const int64_t gbase = block_off[blockIdx.x];
    // LDS byte j holds text[gbase - lead + j], with text + gbase - lead 16-byte aligned
    const int lead = (int)((reinterpret_cast<uintptr_t>(text) + (uintptr_t)gbase) & 15);
    // positions are 32-bit offsets into the block's text (< kStage bytes)
    auto put = [&](int pos, char ch) { stage[lead + pos] = (unsigned char)ch; };
    int pos = (int)off;
    int64_t col = i0 < n ? i0 % row_len : 0;  // row position: ',' after an element, '\n' after a row's last
#pragma unroll
    for (int k = 0; k < kFPer; ++k) {
        const int64_t i = i0 + k;
        if (i < n) {
            if (FLOAT) {
                if (dg[k].flags & 1) special_write(dg[k].N, put, pos);
                else g17_write(g17_make(dg[k].flags & 2, dg[k].N, dg[k].X), put, pos);
            } else {
                i64_write((int64_t)b[k], put, pos);
            }
            put(pos + lens[k] - 1, col == row_len - 1 ? '\n' : ',');
            col = col == row_len - 1 ? 0 : col + 1;
            pos += lens[k];
        }
    }
    __syncthreads();
    // LDS -> HBM: whole 16-byte lines inside [gbase, gbase + total) as one
    // store each, the partial first and last lines byte by byte
    const int64_t end = gbase + total;
    const int64_t a0 = gbase - lead;  // 16-aligned
    const int64_t lines = (end - a0 + 15) >> 4;
    for (int64_t L = threadIdx.x; L < lines; L += kFThreads) {
        const int64_t a = a0 + 16 * L;
        if (a >= gbase && a + 16 <= end) {
            *reinterpret_cast<uint4*>(text + a) = *reinterpret_cast<const uint4*>(stage + 16 * L);
        } else {
            for (int b = 0; b < 16; ++b)
                if (a + b >= gbase && a + b < end) text[a + b] = stage[16 * L + b];
        }
    }
}

int64_t bitmap_words64(int64_t count) { return ((count + 31) / 32 + 1) / 2 + 1; }

// Test mode (hydhip_csv_test_force_exact): every fp64 value is formatted by
// the exact kernel.
int g_force_exact = 0;

}  // namespace

// ---- host side ------------------------------------------------------------
int csv_parse(const unsigned char* text, int64_t nbytes, int64_t count, bool is_float, int64_t* out,
              hydhip_csv_status_t* st, hipStream_t s) {
    st->error_element = -1;
    st->error_kind = HYDHIP_CSV_OK;
    st->exact_path = 0;
    st->numbers_found = 0;
    if (count == 0) return HYD_OK;
    if (nbytes == 0) {
        st->error_element = 0;
        st->error_kind = HYDHIP_CSV_PARSE_ERROR;
        return HYD_OK;
    }
    const int64_t tiles = ceil_div(nbytes, kTile);
    if (tiles > 0x7fffffff) return set_error(HYD_E_INVALID, "csv: text too large");
    const int64_t bw = bitmap_words64(count);
    const int64_t words = 3 * tiles + scan_workspace_words(tiles) + bw + 8;
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)words * 8));
    if (!ws) return set_error(HYD_E_OOM, "csv: workspace of %lld words", (long long)words);
    int64_t* tile_cnt = ws;
    int64_t* tile_err = tile_cnt + tiles;
    int64_t* tile_base = tile_err + tiles;
    int64_t* scan_ws = tile_base + tiles;
    int64_t* status = scan_ws + scan_workspace_words(tiles);       // 4 words
    unsigned long long* exact_cnt = reinterpret_cast<unsigned long long*>(status + 4);
    unsigned long long* err_enc = reinterpret_cast<unsigned long long*>(status + 5);
    uint32_t* exact_bits = reinterpret_cast<uint32_t*>(status + 8);
    unsigned long long* irregular = reinterpret_cast<unsigned long long*>(status + 6);
    const bool aligned = (reinterpret_cast<uintptr_t>(text) & 15) == 0;
    const unsigned g = (unsigned)tiles;
    int64_t h[7];
    // pass 1 assumes one number per word; pass 2 (only when a word broke
    // that) counts the numbers first
    for (int pass = 0; pass < 2; ++pass) {
        HYD_TRY(fill_zero(status, (size_t)(8 + bw) * 8, s));
#define HYD_PARSE_LAUNCH(F, A)                                                                            \
    do {                                                                                                  \
        if (pass == 0) csv_words_kernel<A><<<g, kPThreads, 0, s>>>(text, nbytes, tile_cnt);               \
        else csv_count_kernel<F, A><<<g, kPThreads, 0, s>>>(text, nbytes, tile_cnt, tile_err);            \
        HYD_CHECK_LAUNCH();                                                                               \
        HYD_TRY(scan_i64(tile_cnt, tiles, 0, tile_base, false, false, s, scan_ws));                      \
        if (pass == 0)                                                                                    \
            csv_value_kernel<F, A, true><<<g, kPThreads, 0, s>>>(text, nbytes, tile_base, tile_err, count, \
                                                                 out, exact_bits, exact_cnt, err_enc,     \
                                                                 irregular);                             \
        else                                                                                              \
            csv_value_kernel<F, A, false><<<g, kPThreads, 0, s>>>(text, nbytes, tile_base, tile_err, count, \
                                                                  out, exact_bits, exact_cnt, err_enc,    \
                                                                  irregular);                            \
        HYD_CHECK_LAUNCH();                                                                               \
    } while (0)
        if (is_float) {
            if (aligned) HYD_PARSE_LAUNCH(true, true);
            else HYD_PARSE_LAUNCH(true, false);
            csv_exact_kernel<<<64, 64, 0, s>>>(text, nbytes, count, exact_bits, exact_cnt, out);
            HYD_CHECK_LAUNCH();
        } else {
            if (aligned) HYD_PARSE_LAUNCH(false, true);
            else HYD_PARSE_LAUNCH(false, false);
        }
#undef HYD_PARSE_LAUNCH
        csv_status_kernel<<<1, kPThreads, 0, s>>>(text, nbytes, count, tiles, tile_cnt, tile_base, exact_cnt,
                                                  err_enc, status);
        HYD_CHECK_LAUNCH();
        HYD_TRY_HIP(hipMemcpyAsync(h, status, sizeof h, hipMemcpyDeviceToHost, s));
        HYD_TRY_HIP(hipStreamSynchronize(s));
        if (!h[6]) break;
    }
    st->error_element = h[0];
    st->error_kind = (int32_t)h[1];
    st->numbers_found = h[2];
    st->exact_path = h[3];
    return HYD_OK;
}

int csv_format(const uint64_t* vals, int64_t count, int64_t row_len, bool is_float, unsigned char* text,
               int64_t capacity, int64_t* nbytes, int64_t* exact, hipStream_t s) {
    *nbytes = 0;
    if (exact) *exact = 0;
    if (count == 0) return HYD_OK;
    if (row_len <= 0) return set_error(HYD_E_INVALID, "csv: row length %lld", (long long)row_len);
    const int64_t blocks = ceil_div(count, kFBlock);
    if (blocks > 0x7fffffff) return set_error(HYD_E_INVALID, "csv: too many values");
    const int64_t bw = bitmap_words64(count);
    const int64_t dig_words = is_float ? 3 * count : 0;  // sparse exact digits, packed table digits
    const int64_t words = 2 * blocks + scan_workspace_words(blocks) + 8 + bw + dig_words;
    int64_t* ws = static_cast<int64_t*>(stream_workspace(s, (size_t)words * 8));
    if (!ws) return set_error(HYD_E_OOM, "csv: workspace of %lld words", (long long)words);
    int64_t* block_len = ws;
    int64_t* block_off = block_len + blocks;
    int64_t* scan_ws = block_off + blocks;
    int64_t* misc = scan_ws + scan_workspace_words(blocks);  // [0] exact count
    unsigned long long* exact_cnt = reinterpret_cast<unsigned long long*>(misc);
    uint32_t* exact_bits = reinterpret_cast<uint32_t*>(misc + 8);
    Dig* dig = reinterpret_cast<Dig*>(misc + 8 + bw);
    uint64_t* packed = reinterpret_cast<uint64_t*>(misc + 8 + bw + 2 * count);
    const unsigned g = (unsigned)blocks;
    HYD_TRY(fill_zero(misc, (size_t)(8 + bw) * 8, s));
    if (is_float) {
        if (g_force_exact)
            fmt_len_kernel<true, true><<<g, kFThreads, 0, s>>>(vals, count, block_len, exact_bits, exact_cnt, packed);
        else
            fmt_len_kernel<true, false><<<g, kFThreads, 0, s>>>(vals, count, block_len, exact_bits, exact_cnt, packed);
        HYD_CHECK_LAUNCH();
        fmt_exact_kernel<<<64, 64, 0, s>>>(vals, count, dig, exact_bits, exact_cnt, block_len);
    } else {
        fmt_len_kernel<false, false><<<g, kFThreads, 0, s>>>(vals, count, block_len, exact_bits, exact_cnt,
                                                             nullptr);
    }
    HYD_CHECK_LAUNCH();
    HYD_TRY(scan_i64(block_len, blocks, 0, block_off, false, false, s, scan_ws));
    int64_t h[3];
    HYD_TRY_HIP(hipMemcpyAsync(&h[0], block_off + blocks - 1, 8, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipMemcpyAsync(&h[1], block_len + blocks - 1, 8, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipMemcpyAsync(&h[2], misc, 8, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipStreamSynchronize(s));
    const int64_t total = h[0] + h[1];
    if (exact) *exact = is_float ? h[2] : 0;
    if (total > capacity)
        return set_error(HYD_E_INVALID, "csv: text needs %lld bytes, capacity %lld", (long long)total,
                         (long long)capacity);
    if (is_float)
        fmt_write_kernel<true><<<g, kFThreads, 0, s>>>(vals, count, row_len, block_off, text, dig, exact_bits,
                                                       exact_cnt, packed);
    else
        fmt_write_kernel<false><<<g, kFThreads, 0, s>>>(vals, count, row_len, block_off, text, dig, exact_bits,
                                                        exact_cnt, nullptr);
    HYD_CHECK_LAUNCH();
    *nbytes = total;
    return HYD_OK;
}

// ---- file I/O and large host copies (pinned double buffers) ----------------
// A file's bytes move through two pinned halves: while one half is on PCIe,
// the host fills (read: several threads' pread of disjoint pieces, since one
// thread copying from the page cache is far slower than the link) or drains
// (write: pwrite) the other.
namespace {

constexpr size_t kIoHalf = (size_t)256 << 20;

struct PinnedPair {
    std::mutex mu;
    unsigned char* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
};
PinnedPair g_io;

int io_buffers() {
    for (int k = 0; k < 2; ++k) {
        if (!g_io.host[k]) HYD_TRY_HIP(hipHostMalloc((void**)&g_io.host[k], kIoHalf, hipHostMallocDefault));
        if (!g_io.ev[k]) HYD_TRY_HIP(hipEventCreateWithFlags(&g_io.ev[k], hipEventDisableTiming));
    }
    return HYD_OK;
}

int io_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return hw == 0 ? 4 : (hw > 16 ? 16 : (int)hw);
}

// buf[0, len) <-> file bytes [off, off + len): reads by io_threads()
// threads; writes by one (the kernel serialises writers of one file on its
// inode lock, so more threads only add contention).
// On failure *err is the errno of the failing call, or EIO when a write
// stored nothing (0 returned, errno untouched) or a read met the file's end.
bool parallel_io(int fd, unsigned char* buf, size_t len, off_t off, bool write, int* err) {
    const int nt = (write || len < ((size_t)4 << 20)) ? 1 : io_threads();
    std::atomic<bool> ok{true};
    std::atomic<int> first_err{0};
    auto piece = [&](int t) {
        const size_t a = len * t / nt, e = len * (t + 1) / nt;
        size_t done = a;
        while (done < e && ok.load()) {
            const ssize_t r = write ? pwrite(fd, buf + done, e - done, off + (off_t)done)
                                    : pread(fd, buf + done, e - done, off + (off_t)done);
            if (r > 0) {
                done += (size_t)r;
            } else if (r < 0 && errno == EINTR) {
                continue;
            } else {  // error, or the file ended early
                int expected = 0;
                first_err.compare_exchange_strong(expected, r < 0 ? errno : EIO);
                ok = false;
            }
        }
    };
    if (nt == 1) {
        piece(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(piece, t);
        for (auto& t : th) t.join();
    }
    if (!ok.load() && err) *err = first_err.load();
    return ok.load();
}

// Armed for a call that queues copies into the pinned halves on s: on an
// early (error) return it waits for every copy the call queued, so that no
// later caller gets a half that a copy still reads or fills.
struct DrainOnError {
    hipStream_t s;
    bool ok = false;
    explicit DrainOnError(hipStream_t st) : s(st) {}
    ~DrainOnError() {
        if (ok) return;
        (void)hipStreamSynchronize(s);
        g_io.pending[0] = g_io.pending[1] = false;
    }
};

// dst[0, len) = src[0, len) by io_threads() threads (one below 4 MiB): a
// single thread copying pageable memory into the pinned halves is slower
// than the link.
void parallel_copy(unsigned char* dst, const unsigned char* src, size_t len) {
    const int nt = len < ((size_t)4 << 20) ? 1 : io_threads();
    if (nt == 1) {
        memcpy(dst, src, len);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const size_t a = len * t / nt, e = len * (t + 1) / nt;
        th.emplace_back([=]() { memcpy(dst + a, src + a, e - a); });
    }
    for (auto& t : th) t.join();
}

[[noreturn]] void die_hip(const char* what) {
    fprintf(stderr, "%s: %s\n", what, hydhip_last_error());
    exit(1);
}

// file bytes [0, nbytes) -> dev; the copies are queued on s.
int file_to_device(int fd, unsigned char* dev, int64_t nbytes, hipStream_t s, bool* short_read) {
    std::lock_guard<std::mutex> lk(g_io.mu);
    HYD_TRY(io_buffers());
    DrainOnError drain(s);
    *short_read = false;
    int64_t c = 0;
    for (int64_t off = 0; off < nbytes; off += (int64_t)kIoHalf, ++c) {
        const int k = (int)(c & 1);
        const size_t len = (size_t)std::min<int64_t>((int64_t)kIoHalf, nbytes - off);
        if (g_io.pending[k]) HYD_TRY_HIP(hipEventSynchronize(g_io.ev[k]));
        g_io.pending[k] = false;
        if (!parallel_io(fd, g_io.host[k], len, (off_t)off, false, nullptr)) {
            *short_read = true;
            break;
        }
        HYD_TRY_HIP(hipMemcpyAsync(dev + off, g_io.host[k], len, hipMemcpyHostToDevice, s));
        HYD_TRY_HIP(hipEventRecord(g_io.ev[k], s));
        g_io.pending[k] = true;
    }
    for (int k = 0; k < 2; ++k)
        if (g_io.pending[k]) {
            HYD_TRY_HIP(hipEventSynchronize(g_io.ev[k]));
            g_io.pending[k] = false;
        }
    drain.ok = true;
    return HYD_OK;
}

// dev[0, bytes) -> file bytes [0, bytes): the copy of half i + 1 overlaps the
// pwrite of half i.
int device_to_file(int fd, const unsigned char* dev, int64_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_io.mu);
    HYD_TRY(io_buffers());
    DrainOnError drain(s);
    const int64_t halves = ceil_div(bytes, (int64_t)kIoHalf);
    auto issue = [&](int64_t c) -> int {
        const int64_t off = c * (int64_t)kIoHalf;
        const int64_t len = std::min<int64_t>((int64_t)kIoHalf, bytes - off);
        HYD_TRY_HIP(hipMemcpyAsync(g_io.host[c & 1], dev + off, (size_t)len, hipMemcpyDeviceToHost, s));
        HYD_TRY_HIP(hipEventRecord(g_io.ev[c & 1], s));
        return HYD_OK;
    };
    if (halves > 0) HYD_TRY(issue(0));
    for (int64_t c = 0; c < halves; ++c) {
        if (c + 1 < halves) HYD_TRY(issue(c + 1));  // its buffer's pwrite (half c - 1) has finished
        HYD_TRY_HIP(hipEventSynchronize(g_io.ev[c & 1]));
        const int64_t off = c * (int64_t)kIoHalf;
        const int64_t len = std::min<int64_t>((int64_t)kIoHalf, bytes - off);
        int err = 0;
        if (!parallel_io(fd, g_io.host[c & 1], (size_t)len, (off_t)off, true, &err))
            return set_error(HYD_E_INVALID, "csv: write failed: %s", strerror(err));
    }
    drain.ok = true;
    return HYD_OK;
}

}  // namespace

// Host memory <-> HBM through the pinned halves: the host copy of one half
// (parallel memcpy) overlaps the DMA of the other.  Both return when the
// data has arrived.  For large pageable buffers, where hipMemcpy would stage
// through the runtime's own small buffers one copy at a time.
int staged_h2d(void* dev, const void* host, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_io.mu);
    HYD_TRY(io_buffers());
    DrainOnError drain(s);
    unsigned char* d = static_cast<unsigned char*>(dev);
    const unsigned char* h = static_cast<const unsigned char*>(host);
    int64_t c = 0;
    for (size_t off = 0; off < bytes; off += kIoHalf, ++c) {
        const int k = (int)(c & 1);
        const size_t len = std::min(kIoHalf, bytes - off);
        if (g_io.pending[k]) HYD_TRY_HIP(hipEventSynchronize(g_io.ev[k]));
        g_io.pending[k] = false;
        parallel_copy(g_io.host[k], h + off, len);
        HYD_TRY_HIP(hipMemcpyAsync(d + off, g_io.host[k], len, hipMemcpyHostToDevice, s));
        HYD_TRY_HIP(hipEventRecord(g_io.ev[k], s));
        g_io.pending[k] = true;
    }
    for (int k = 0; k < 2; ++k)
        if (g_io.pending[k]) {
            HYD_TRY_HIP(hipEventSynchronize(g_io.ev[k]));
            g_io.pending[k] = false;
        }
    drain.ok = true;
    return HYD_OK;
}

int staged_d2h(void* host, const void* dev, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_io.mu);
    HYD_TRY(io_buffers());
    DrainOnError drain(s);
    unsigned char* h = static_cast<unsigned char*>(host);
    const unsigned char* d = static_cast<const unsigned char*>(dev);
    const int64_t halves = ceil_div((int64_t)bytes, (int64_t)kIoHalf);
    auto issue = [&](int64_t c) -> int {
        const size_t off = (size_t)c * kIoHalf;
        const size_t len = std::min(kIoHalf, bytes - off);
        HYD_TRY_HIP(hipMemcpyAsync(g_io.host[c & 1], d + off, len, hipMemcpyDeviceToHost, s));
        HYD_TRY_HIP(hipEventRecord(g_io.ev[c & 1], s));
        return HYD_OK;
    };
    if (halves > 0) HYD_TRY(issue(0));
    for (int64_t c = 0; c < halves; ++c) {
        if (c + 1 < halves) HYD_TRY(issue(c + 1));  // its buffer's host copy (half c - 1) has finished
        HYD_TRY_HIP(hipEventSynchronize(g_io.ev[c & 1]));
        const size_t off = (size_t)c * kIoHalf;
        parallel_copy(h + off, g_io.host[c & 1], std::min(kIoHalf, bytes - off));
    }
    drain.ok = true;
    return HYD_OK;
}

namespace {

// Device copy of a file's bytes (a buffer kept for the next call).
struct TextBuffer {
    std::mutex mu;
    unsigned char* dev = nullptr;
    size_t bytes = 0;
};
TextBuffer g_text;

int text_buffer(size_t bytes, unsigned char** out) {
    if (g_text.bytes < bytes) {
        if (g_text.dev) HYD_TRY_HIP(hipFree(g_text.dev));
        g_text.dev = nullptr;
        g_text.bytes = 0;
        const size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
        HYD_TRY_HIP(hipMalloc(&g_text.dev, want));
        g_text.bytes = want;
    }
    *out = g_text.dev;
    return HYD_OK;
}

}  // namespace

// Frees the CSV file buffers (pinned halves, HBM text buffer); from the
// library's cleanup.
void release_csv_buffers() {
    {
        std::lock_guard<std::mutex> lk(g_io.mu);
        for (int k = 0; k < 2; ++k) {
            if (g_io.pending[k] && g_io.ev[k]) (void)hipEventSynchronize(g_io.ev[k]);
            if (g_io.ev[k]) (void)hipEventDestroy(g_io.ev[k]);
            if (g_io.host[k]) (void)hipHostFree(g_io.host[k]);
            g_io.ev[k] = nullptr;
            g_io.host[k] = nullptr;
            g_io.pending[k] = false;
        }
    }
    std::lock_guard<std::mutex> lk(g_text.mu);
    if (g_text.dev) (void)hipFree(g_text.dev);
    g_text.dev = nullptr;
    g_text.bytes = 0;
}

namespace {

hyd_array_t* read_csv_file(const char* filename, hyd_tuple_t shape, bool is_float, const char* who) {
    if (ensure_init() != HYD_OK) die_hip(who);
    hyd_array_t* arr = hyd_array_alloc_device(shape, sizeof(int64_t));
    const int64_t count = hyd_shape_size(shape);
    FILE* f = fopen(filename, "r");
    if (!f) {
        fprintf(stderr, "%s: cannot open '%s'\n", who, filename);
        exit(1);
    }
    hipStream_t s = export_stream();
    std::lock_guard<std::mutex> lk(g_text.mu);
    unsigned char* dev = nullptr;
    int64_t nbytes = 0;
    struct stat sb;
    const int fd = fileno(f);
    if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode)) {
        nbytes = (int64_t)sb.st_size;
        if (text_buffer((size_t)nbytes + 16, &dev) != HYD_OK) die_hip(who);
        bool short_read = false;
        if (file_to_device(fd, dev, nbytes, s, &short_read) != HYD_OK) die_hip(who);
        if (short_read) {
            fprintf(stderr, "%s: cannot read '%s'\n", who, filename);
            exit(1);
        }
    } else {
        // a pipe or similar: read it all, then copy
        std::string all;
        char tmp[1 << 16];
        size_t got;
        while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) all.append(tmp, got);
        nbytes = (int64_t)all.size();
        if (text_buffer((size_t)nbytes + 16, &dev) != HYD_OK) die_hip(who);
        if (nbytes && (hipMemcpyAsync(dev, all.data(), all.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
                       hipStreamSynchronize(s) != hipSuccess))
            die_hip(who);
    }
    fclose(f);
    hydhip_csv_status_t st;
    if (csv_parse(dev, nbytes, count, is_float, static_cast<int64_t*>(hyd_array_device_data(arr)), &st, s) !=
        HYD_OK)
        die_hip(who);
    if (st.error_kind == HYDHIP_CSV_TOO_FEW) {
        fprintf(stderr, "%s: too few values\n", who);
        exit(1);
    }
    if (st.error_kind == HYDHIP_CSV_PARSE_ERROR) {
        fprintf(stderr, "%s: parse error at element %lld\n", who, (long long)st.error_element);
        exit(1);
    }
    mark_device_written(arr);
    if (host_sync_enabled() && hyd_array_sync_host(arr) != HYD_OK) die_hip(who);
    return arr;
}

int64_t write_csv_file(hyd_array_t* arr, const char* filename, bool is_float) {
    FILE* f = fopen(filename, "w");
    if (!f) {
        perror("fopen");
        exit(1);
    }
    const int ndims = arr->shape.ndims;
    const int64_t size = hyd_shape_size(arr->shape);
    const int64_t rows = ndims == 2 ? arr->shape.elems[0] : 1;
    const int64_t row_len = ndims == 2 ? arr->shape.elems[1] : size;
    if (size == 0) {
        for (int64_t i = 0; i < rows; ++i) fputc('\n', f);  // empty rows
    } else {
        if (ensure_init() != HYD_OK) die_hip("write_array");
        hipStream_t s = export_stream();
        const uint64_t* vals = static_cast<const uint64_t*>(hyd_array_device_data(arr));
        std::lock_guard<std::mutex> lk(g_text.mu);
        const int64_t cap = size * kMaxText;
        unsigned char* dev = nullptr;
        int64_t nbytes = 0;
        if (text_buffer((size_t)cap, &dev) != HYD_OK ||
            csv_format(vals, size, row_len, is_float, dev, cap, &nbytes, nullptr, s) != HYD_OK ||
            device_to_file(fileno(f), dev, nbytes, s) != HYD_OK)
            die_hip("write_array");
    }
    fclose(f);
    return 0;
}

}  // namespace
}  // namespace hyd

// ---- C ABI ----------------------------------------------------------------
using namespace hyd;

extern "C" int hydhip_csv_parse_i64(const void* text, int64_t nbytes, int64_t count, int64_t* out,
                                    hydhip_csv_status_t* status, void* stream) {
    clear_error();
    HYD_TRY(ensure_init());
    if (nbytes < 0 || count < 0 || !status || (count && !out) || (nbytes && !text))
        return set_error(HYD_E_INVALID, "csv_parse_i64: bad arguments");
    return csv_parse(static_cast<const unsigned char*>(text), nbytes, count, false, out, status,
                     resolve_stream(stream));
}

extern "C" int hydhip_csv_parse_f64(const void* text, int64_t nbytes, int64_t count, double* out,
                                    hydhip_csv_status_t* status, void* stream) {
    clear_error();
    HYD_TRY(ensure_init());
    if (nbytes < 0 || count < 0 || !status || (count && !out) || (nbytes && !text))
        return set_error(HYD_E_INVALID, "csv_parse_f64: bad arguments");
    return csv_parse(static_cast<const unsigned char*>(text), nbytes, count, true, reinterpret_cast<int64_t*>(out),
                     status, resolve_stream(stream));
}

extern "C" int64_t hydhip_csv_text_bound(int64_t count) { return count < 0 ? 0 : count * kMaxText; }

extern "C" int hydhip_csv_format_i64(const int64_t* vals, int64_t count, int64_t row_len, void* text,
                                     int64_t capacity, int64_t* nbytes, void* stream) {
    clear_error();
    HYD_TRY(ensure_init());
    if (count < 0 || !nbytes || (count && (!vals || !text)))
        return set_error(HYD_E_INVALID, "csv_format_i64: bad arguments");
    return csv_format(reinterpret_cast<const uint64_t*>(vals), count, row_len, false,
                      static_cast<unsigned char*>(text), capacity, nbytes, nullptr, resolve_stream(stream));
}

extern "C" int hydhip_csv_format_f64(const double* vals, int64_t count, int64_t row_len, void* text,
                                     int64_t capacity, int64_t* nbytes, int64_t* exact_path, void* stream) {
    clear_error();
    HYD_TRY(ensure_init());
    if (count < 0 || !nbytes || (count && (!vals || !text)))
        return set_error(HYD_E_INVALID, "csv_format_f64: bad arguments");
    return csv_format(reinterpret_cast<const uint64_t*>(vals), count, row_len, true, static_cast<unsigned char*>(text),
                      capacity, nbytes, exact_path, resolve_stream(stream));
}

// Drop-in runtime functions (runtime/hydrangea_runtime.h:549-594 and
// runtime/hyd_write_csv.c:24-59): same names, arguments, messages and exit(1).
extern "C" hyd_array_t* hyd_read_array_csv(const char* filename, hyd_tuple_t shape) {
    return read_csv_file(filename, shape, false, "read_array");
}

extern "C" hyd_array_t* hyd_read_float_array_csv(const char* filename, hyd_tuple_t shape) {
    return read_csv_file(filename, shape, true, "read_array_float");
}

extern "C" int64_t hyd_write_array_csv(hyd_array_t* arr, const char* filename) {
    return write_csv_file(arr, filename, false);
}

extern "C" int64_t hyd_write_array_csv_float(hyd_array_t* arr, const char* filename) {
    return write_csv_file(arr, filename, true);
}

// CPU test hook: per-byte gap flags of bytes[0, n) by the 8-byte classifier the
// word kernels use (n a multiple of 8).
extern "C" void hydhip_host_csv_gap_flags(const void* bytes, int64_t n, unsigned char* out) {
    for (int64_t i = 0; i + 8 <= n; i += 8) {
        uint64_t x;
        memcpy(&x, static_cast<const unsigned char*>(bytes) + i, 8);
        const uint32_t m = gap_mask8(x);
        for (int k = 0; k < 8; ++k) out[i + k] = (m >> k) & 1;
    }
}

// Test hook: 1 = format every fp64 value through the exact kernel (the path
// the table leaves undecided values to, which real values almost never take).
extern "C" void hydhip_csv_test_force_exact(int on) { g_force_exact = on ? 1 : 0; }

// ---- CPU test hooks: the same conversions, compiled for the host ------------
// fscanf-like: skips isspace, then reads one number from s[0, len).  Returns 1
// (value in *value, characters used in *consumed), 0 (no number) or -1 (only
// whitespace); *exact = 1 when the big-integer path rounded it.  flags & 1
// sends every decimal through the big-integer path (to test that path).
extern "C" int hydhip_host_csv_scan(const char* s, int64_t len, int is_float, int flags, void* value,
                                    int64_t* consumed, int* exact) {
    auto at = [&](int64_t p) -> int { return (p < 0 || p >= len) ? -1 : (unsigned char)s[p]; };
    int64_t p = 0;
    while (p < len && is_ws(at(p))) ++p;
    *exact = 0;
    if (p >= len) return -1;
    if (!is_float) {
        const int64_t l = scan_i64(at, p, static_cast<int64_t*>(value));
        if (l < 0) return 0;
        *consumed = p + l;
        return 1;
    }
    FloatTok t;
    bool swar = false;
    if (flags & 2) {  // the 8-byte path, over the text followed by gap bytes as a file would have
        std::string padded(s + p, (size_t)(len - p));
        padded.append(kSwarSpan + 16, ',');
        auto atp = [&](int64_t q) -> int { return (q < 0 || q >= (int64_t)padded.size()) ? -1 : (unsigned char)padded[q]; };
        auto rd8 = [&](int64_t q) -> uint64_t {
            uint64_t v;
            memcpy(&v, padded.data() + q, 8);
            return v;
        };
        swar = scan_dec_swar(rd8, atp, 0, &t);
        if (swar) *exact = 2;
    }
    if (!swar) t = scan_f64_fast(at, p);
    if (t.len < 0) return 0;
    uint64_t bits;
    const bool decided = tok_to_bits(t, &bits);
    if (!decided || ((flags & 1) && t.kind == kDec && t.w != 0)) {
        Big* big = new Big[3];
        bits = dec_exact_bits(at, p, bits, big[0], big[1], big[2]);
        delete[] big;
        *exact |= 1;
    }
    memcpy(value, &bits, 8);
    *consumed = p + t.len;
    return 1;
}

// "%.17g" / "%lld" text of each value (no separator) at out + k * stride;
// returns how many fp64 values took the big-integer path.  flags & 1 sends
// every finite nonzero fp64 value through it.
extern "C" int64_t hydhip_host_csv_format(const void* vals, int64_t n, int is_float, int flags, char* out,
                                          int64_t stride, int32_t* lens) {
    int64_t exact = 0;
    Big* big = new Big[3];
    for (int64_t k = 0; k < n; ++k) {
        char* o = out + k * stride;
        auto put = [&](int64_t pos, char ch) { o[pos] = ch; };
        uint64_t b;
        memcpy(&b, static_cast<const char*>(vals) + 8 * k, 8);
        if (!is_float) {
            lens[k] = i64_write((int64_t)b, put, 0);
            continue;
        }
        if (const int sl = special_len(b)) {
            special_write(b, put, 0);
            lens[k] = sl;
            continue;
        }
        uint64_t N;
        int X;
        const bool decided = (flags & 1) ? f64_to_dec17<true>(b & ~kSignBit, &N, &X)
                                         : f64_to_dec17(b & ~kSignBit, &N, &X);
        if (!decided) {
            dec17_exact(b & ~kSignBit, N, X, &N, &X, big[0], big[1], big[2]);
            ++exact;
        }
        lens[k] = g17_write(g17_make(b >> 63, N, X), put, 0);
    }
    delete[] big;
    return exact;
}
