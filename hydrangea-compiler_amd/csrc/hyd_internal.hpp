// hyd_internal.hpp - shared host/device helpers of libhydhip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/hyd_hip.h"

namespace hyd {

// Records a formatted message for hydhip_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Makes sure the runtime is initialised (device selected, library stream
// created).  Returns HYD_OK or an error code.
int ensure_init();

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

// Zero-fill kernel launch (graph-capture safe, unlike hipMemsetAsync).
int fill_zero(void* dptr, size_t bytes, hipStream_t stream);

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

// C's truncating remainder x % m (CodegenC.hs:2473-2474) for m >= 1, fast:
// a mask when m is a power of two (pow2 = m - 1, else 0), 32-bit division
// when x and m fit in 32 bits, the 64-bit division otherwise.  Negative x
// gives the C result (in (-m, 0]).
__device__ __forceinline__ int64_t c_mod(int64_t x, int64_t m, int64_t pow2) {
    if (pow2) {
        const int64_t r = x & pow2;
        return (x < 0 && r) ? r - m : r;
    }
    if ((uint64_t)(x + 0x80000000ll) < 0x100000000ull && m <= 0x7fffffff)
        return (int64_t)((int32_t)x % (int32_t)m);
    return x % m;
}
inline int64_t pow2_mask(int64_t m) { return (m > 0 && (m & (m - 1)) == 0) ? m - 1 : 0; }

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
