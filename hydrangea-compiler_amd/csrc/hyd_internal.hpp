// hyd_internal.hpp - shared host/device helpers of libhydhip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <functional>

#include "../../include/hyd_hip.h"

namespace hyd {

// Records a formatted message for hydhip_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Makes sure the runtime is initialised (device selected, library stream
// created).  Returns HYD_OK or an error code.
int ensure_init();

// Whether ensure_init() has succeeded (no side effects).
bool library_initialized();

// Looks up, once per device, the device address of the no-allocation
// histogram's file-scope scratch array (scatter.hip); `device` must be the
// current device.  Called by ensure_init() for the library's device.
int noalloc_scratch_setup(int device);

// NULL -> the library's stream.
hipStream_t resolve_stream(void* stream);

struct DeviceInfo {
    int device = 0;
    int num_cus = 0;
    size_t lds_per_block = 0;  // bytes usable by one workgroup
    size_t hbm_bytes = 0;
    char name[128] = {0};
};
const DeviceInfo& device_info();

// Scratch memory tied to a stream: a device buffer of at least `bytes`,
// reused by later calls on the same stream (kernels on one stream run in
// order, so the buffer is free again when the next call's kernels start).
// Growing it frees the old buffer after a device synchronisation.
void* stream_workspace(hipStream_t stream, size_t bytes);

// A small device buffer for scalar results of the hyd_export_* wrappers
// (which, like the reference's, are not reentrant).
void* scalar_slot();

// A per-stream block of device memory that is all zero between kernels:
// kernels that use it (the spread merge of the scatters) clear what they
// touched before they finish.  kZeroScratchBytes long, 256-byte aligned.
constexpr size_t kZeroScratchBytes = (size_t)1 << 20;
void* stream_zero_scratch(hipStream_t stream);

// Copies `bytes` of descriptors, written by fill() into a pinned host buffer,
// to dev on stream s (asynchronously; the buffer is reused only once the copy
// that read it has completed).  Not for use under graph capture.
int stage_h2d(hipStream_t s, void* dev, size_t bytes, const std::function<void(unsigned char*)>& fill);

// Zero-fill kernel launch (graph-capture safe, unlike hipMemsetAsync).
int fill_zero(void* dptr, size_t bytes, hipStream_t stream);

// Large pageable host buffer <-> HBM through two pinned 256 MiB halves, the
// host copy of one half (several threads) overlapping the DMA of the other
// (csv.hip).  Both return when the data has arrived.
int staged_h2d(void* dev, const void* host, size_t bytes, hipStream_t s);
int staged_d2h(void* host, const void* dev, size_t bytes, hipStream_t s);

// The batched-scatter autotuner's table as a dense array, mode[ceil log2
// bins][floor log2 n]: 0 = LDS tiles, 1 = global atomics, -1 = untuned.
struct TuneSnapshot {
    int8_t mode[64][64];
};

// The tuned streaming grid (workgroups per CU) of a launch-shape kind
// (HYD_SHAPE_*) for n elements, or dflt when that class is untuned.
int tuned_per_cu(int kind, int64_t n, int dflt);

}  // namespace hyd

#define HYD_TRY_HIP(expr)                                                          \
    do {                                                                           \
        hipError_t hyd_e_ = (expr);                                                \
        if (hyd_e_ != hipSuccess)                                                  \
            return hyd::set_error(HYD_E_HIP, "%s failed: %s", #expr,               \
                                  hipGetErrorString(hyd_e_));                      \
    } while (0)

#define HYD_TRY(expr)                   \
    do {                                \
        int hyd_rc_ = (expr);           \
        if (hyd_rc_ != HYD_OK) return hyd_rc_; \
    } while (0)

// After a kernel launch: report launch-configuration errors.
#define HYD_CHECK_LAUNCH() HYD_TRY_HIP(hipGetLastError())

// ---- device helpers ------------------------------------------------------
namespace hyd {

__host__ __device__ inline uint64_t splitmix_at(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// C's truncating remainder x % m (CodegenC.hs:2473-2474) for an invariant
// divisor m >= 1, without a division or a branch: unsigned division of |x| by
// multiplication with a precomputed magic number (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1:
// l = ceil(log2 m), magic = floor(2^64 (2^l - m) / m) + 1, sh1 = min(l, 1),
// sh2 = max(l - 1, 0); q = (t + ((n - t) >> sh1)) >> sh2 with t = mulhi(magic, n),
// exact for every 0 <= n < 2^64), then the sign of x (result in (-m, m)).
// A branch per element would keep the compiler from issuing the loads of a
// streaming loop ahead of the arithmetic, so callers pick the form once per
// kernel or tile: c_mod_pow2 (a mask) when `mask` is set (m a power of two,
// m >= 2), else c_mod.
struct ModConst {
    uint64_t m;
    uint64_t magic;
    uint32_t sh1, sh2;
    uint64_t mask;  // m - 1 when m >= 2 is a power of two, else 0
};

inline ModConst make_mod_const(int64_t m) {
    ModConst k;
    k.m = (uint64_t)m;
    k.mask = (m >= 2 && (m & (m - 1)) == 0) ? (uint64_t)m - 1 : 0;
    uint32_t l = 0;
    while (l < 64 && ((uint64_t)1 << l) < k.m) ++l;
    const unsigned __int128 num = ((unsigned __int128)(((unsigned __int128)1 << l) - k.m)) << 64;
    k.magic = (uint64_t)(num / k.m) + 1;
    k.sh1 = l < 1 ? l : 1;
    k.sh2 = l > 0 ? l - 1 : 0;
    return k;
}

__host__ __device__ __forceinline__ int64_t c_mod(int64_t x, const ModConst& k) {
    const uint64_t n = x < 0 ? 0 - (uint64_t)x : (uint64_t)x;
    const uint64_t t = (uint64_t)(((unsigned __int128)k.magic * n) >> 64);
    const uint64_t q = (t + ((n - t) >> k.sh1)) >> k.sh2;
    const uint64_t r = n - q * k.m;
    return x < 0 ? -(int64_t)r : (int64_t)r;
}

// x % m for m = k.mask + 1 a power of two (two's complement: the low bits,
// moved into (-m, 0] for a negative x with a nonzero remainder).
__host__ __device__ __forceinline__ int64_t c_mod_pow2(int64_t x, const ModConst& k) {
    const int64_t r = x & (int64_t)k.mask;
    return (x < 0 && r) ? r - (int64_t)k.m : r;
}

// 16-byte streaming accesses (non-temporal: read or written once).
typedef double v2d __attribute__((ext_vector_type(2)));
typedef long long v2i __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v2d ld_stream2(const double* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
}
__device__ __forceinline__ void st_stream2(double* p, v2d v) {
    __builtin_nontemporal_store(v, reinterpret_cast<v2d*>(p));
}
__device__ __forceinline__ v2i ld_stream2(const int64_t* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const v2i*>(p));
}

}  // namespace hyd
