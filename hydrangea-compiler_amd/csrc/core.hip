// core.hip - lifecycle, errors, memory/stream/event plumbing and the
// synthetic-input generators of libhydhip.so.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <functional>
#include <mutex>
#include <unordered_map>

#include "hyd_internal.hpp"

namespace hyd {

static thread_local char g_err[512];

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = 0; }

static std::mutex g_init_mu;
static bool g_inited = false;
static hipStream_t g_stream = nullptr;
static DeviceInfo g_info;

int ensure_init() {
    if (g_inited) return HYD_OK;
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_inited) return HYD_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_error(HYD_E_NODEV, "no HIP device available (%s)",
                         e == hipSuccess ? "count 0" : hipGetErrorString(e));
    int dev = 0;
    HYD_TRY_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HYD_TRY_HIP(hipGetDeviceProperties(&prop, dev));
    g_info.device = dev;
    g_info.num_cus = prop.multiProcessorCount;
    g_info.lds_per_block = prop.sharedMemPerBlock;
    g_info.hbm_bytes = prop.totalGlobalMem;
    snprintf(g_info.name, sizeof g_info.name, "%s", prop.gcnArchName);
    // Default priority only: streams of mixed priorities are not used.
    HYD_TRY_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    HYD_TRY(noalloc_scratch_setup(dev));
    g_inited = true;
    return HYD_OK;
}

bool library_initialized() {
    std::lock_guard<std::mutex> lk(g_init_mu);
    return g_inited;
}

hipStream_t resolve_stream(void* stream) {
    return stream ? static_cast<hipStream_t>(stream) : g_stream;
}

const DeviceInfo& device_info() { return g_info; }

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
static std::mutex g_ws_mu;
static std::unordered_map<hipStream_t, Workspace> g_ws;

void* stream_workspace(hipStream_t stream, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace& w = g_ws[stream];
    if (w.bytes >= bytes && w.ptr) return w.ptr;
    if (w.ptr) {
        // Work still queued on this stream may read the old buffer.
        if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
        (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) return nullptr;
    w.ptr = p;
    w.bytes = want;
    return p;
}

static std::unordered_map<hipStream_t, void*> g_zero;

void* stream_zero_scratch(hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_zero.find(stream);
    if (it != g_zero.end()) return it->second;
    void* p = nullptr;
    if (hipMalloc(&p, kZeroScratchBytes) != hipSuccess) return nullptr;
    // zeroed once; every user leaves it zero again
    if (fill_zero(p, kZeroScratchBytes, stream) != HYD_OK ||
        hipStreamSynchronize(stream) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    g_zero[stream] = p;
    return p;
}

// Per (stream, slot) buffers for generated launchers (hydhip_scratch): kept
// apart from stream_workspace, which the library's own calls use on the same
// stream in between.
static std::unordered_map<hipStream_t, Workspace> g_slots[4];

void* stream_slot(hipStream_t stream, int slot, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace& w = g_slots[slot][stream];
    if (w.bytes >= bytes && w.ptr) return w.ptr;
    if (w.ptr) {
        if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
        (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    size_t want = bytes < (1u << 16) ? (1u << 16) : bytes;
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) return nullptr;
    w.ptr = p;
    w.bytes = want;
    return p;
}

static void* g_scalar = nullptr;

void* scalar_slot() {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (!g_scalar && hipMalloc(&g_scalar, 4096) != hipSuccess) {
        g_scalar = nullptr;
        set_error(HYD_E_OOM, "scalar slot allocation failed");
    }
    return g_scalar;
}

// Pinned staging for host-to-device copies of launch descriptors: two pinned
// buffers per stream, used alternately.  A buffer is refilled only after the
// copy that last read it has completed (an event recorded right after that
// copy), so the host can plan the next call while the GPU still runs this one
// and no call has to wait for its own stream to drain.
namespace {
struct Staging {
    std::mutex mu;
    void* host[2] = {nullptr, nullptr};
    size_t bytes[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    int next = 0;
};
std::mutex g_stage_mu;
std::unordered_map<hipStream_t, Staging*> g_stage;
}  // namespace

int stage_h2d(hipStream_t s, void* dev, size_t bytes, const std::function<void(unsigned char*)>& fill) {
    Staging* st;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        Staging*& slot = g_stage[s];
        if (!slot) slot = new Staging();
        st = slot;
    }
    std::lock_guard<std::mutex> lk(st->mu);
    const int k = st->next;
    st->next ^= 1;
    if (st->pending[k]) {
        HYD_TRY_HIP(hipEventSynchronize(st->ev[k]));
        st->pending[k] = false;
    }
    if (st->bytes[k] < bytes) {
        if (st->host[k]) HYD_TRY_HIP(hipHostFree(st->host[k]));
        st->host[k] = nullptr;
        st->bytes[k] = 0;
        const size_t want = bytes < (64u << 10) ? (64u << 10) : bytes;
        HYD_TRY_HIP(hipHostMalloc(&st->host[k], want, hipHostMallocDefault));
        st->bytes[k] = want;
    }
    if (!st->ev[k]) HYD_TRY_HIP(hipEventCreateWithFlags(&st->ev[k], hipEventDisableTiming));
    fill(static_cast<unsigned char*>(st->host[k]));
    HYD_TRY_HIP(hipMemcpyAsync(dev, st->host[k], bytes, hipMemcpyHostToDevice, s));
    HYD_TRY_HIP(hipEventRecord(st->ev[k], s));
    st->pending[k] = true;
    return HYD_OK;
}

static void release_staging() {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    for (auto& kv : g_stage) {
        Staging* st = kv.second;
        for (int k = 0; k < 2; ++k) {
            if (st->ev[k]) (void)hipEventDestroy(st->ev[k]);
            if (st->host[k]) (void)hipHostFree(st->host[k]);
        }
        delete st;
    }
    g_stage.clear();
}

void release_csv_buffers();  // csv.hip

static void release_workspaces() {
    release_staging();
    release_csv_buffers();
    std::lock_guard<std::mutex> lk(g_ws_mu);
    for (auto& kv : g_ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    g_ws.clear();
    for (auto& kv : g_zero) (void)hipFree(kv.second);
    g_zero.clear();
    if (g_scalar) (void)hipFree(g_scalar);
    g_scalar = nullptr;
}

__global__ void fill_zero_kernel(uint64_t* __restrict__ p, int64_t nwords) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += step) p[i] = 0;
}

__global__ void fill_zero_bytes_kernel(unsigned char* __restrict__ p, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int fill_zero(void* dptr, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return HYD_OK;
    uintptr_t a = reinterpret_cast<uintptr_t>(dptr);
    if ((a & 7) == 0 && (bytes & 7) == 0) {
        int64_t words = (int64_t)(bytes / 8);
        int64_t blocks = ceil_div(words, 256);
        if (blocks > 4096) blocks = 4096;
        fill_zero_kernel<<<(unsigned)blocks, 256, 0, stream>>>(static_cast<uint64_t*>(dptr), words);
    } else {
        int64_t blocks = ceil_div((int64_t)bytes, 256);
        fill_zero_bytes_kernel<<<(unsigned)blocks, 256, 0, stream>>>(static_cast<unsigned char*>(dptr),
                                                                     (int64_t)bytes);
    }
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

__global__ void fill_splitmix_i64_kernel(int64_t* __restrict__ dst, int64_t n, uint64_t seed,
                                         uint64_t c0, int shift) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += step) dst[i] = (int64_t)(splitmix_at(seed, c0 + (uint64_t)i) >> shift);
}

__global__ void fill_splitmix_f64_kernel(double* __restrict__ dst, int64_t n, uint64_t seed,
                                         uint64_t c0, double scale, double offset) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += step) {
        double u = (double)(splitmix_at(seed, c0 + (uint64_t)i) >> 11) * 0x1.0p-53;
        double v = u * scale;
        dst[i] = v + offset;
    }
}

}  // namespace hyd

using namespace hyd;

extern "C" {

const char* hydhip_last_error(void) { return g_err; }

int hyd_hip_init(const char* code_object_dir_or_null) {
    (void)code_object_dir_or_null;  // kernels are linked into the library
    clear_error();
    return ensure_init();
}

void hyd_hip_cleanup(void) {
    if (!g_inited) return;
    (void)hipDeviceSynchronize();
    release_workspaces();
}

int hydhip_device_info(int* num_cus, int64_t* hbm_bytes, int64_t* lds_bytes, char* name,
                       int name_len) {
    HYD_TRY(ensure_init());
    if (num_cus) *num_cus = g_info.num_cus;
    if (hbm_bytes) *hbm_bytes = (int64_t)g_info.hbm_bytes;
    if (lds_bytes) *lds_bytes = (int64_t)g_info.lds_per_block;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", g_info.name);
    return HYD_OK;
}

int hydhip_malloc(void** dptr, int64_t bytes) {
    if (!dptr || bytes < 0) return set_error(HYD_E_INVALID, "hydhip_malloc: bad arguments");
    HYD_TRY(ensure_init());
    *dptr = nullptr;
    if (bytes == 0) return HYD_OK;
    hipError_t e = hipMalloc(dptr, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(HYD_E_OOM, "hipMalloc(%lld) failed: %s", (long long)bytes,
                         hipGetErrorString(e));
    }
    return HYD_OK;
}

int hydhip_free(void* dptr) {
    if (!dptr) return HYD_OK;
    // Kernels queued on any stream may still read the buffer: drain them
    // before the memory goes back to the allocator.
    HYD_TRY_HIP(hipDeviceSynchronize());
    HYD_TRY_HIP(hipFree(dptr));
    return HYD_OK;
}

int hydhip_memcpy_h2d(void* dst, const void* src, int64_t bytes, void* stream) {
    HYD_TRY(ensure_init());
    if (bytes <= 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    HYD_TRY_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, s));
    HYD_TRY_HIP(hipStreamSynchronize(s));
    return HYD_OK;
}

int hydhip_memcpy_d2h(void* dst, const void* src, int64_t bytes, void* stream) {
    HYD_TRY(ensure_init());
    if (bytes <= 0) return HYD_OK;
    hipStream_t s = resolve_stream(stream);
    HYD_TRY_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, s));
    HYD_TRY_HIP(hipStreamSynchronize(s));
    return HYD_OK;
}

int hydhip_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream) {
    HYD_TRY(ensure_init());
    if (bytes <= 0) return HYD_OK;
    HYD_TRY_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice,
                               resolve_stream(stream)));
    return HYD_OK;
}

int hydhip_stream_create(void** stream) {
    HYD_TRY(ensure_init());
    hipStream_t s = nullptr;
    HYD_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return HYD_OK;
}

int hydhip_stream_destroy(void* stream) {
    if (!stream) return HYD_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    (void)hipStreamSynchronize(s);
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto it = g_ws.find(s);
        if (it != g_ws.end()) {
            if (it->second.ptr) (void)hipFree(it->second.ptr);
            g_ws.erase(it);
        }
        auto z = g_zero.find(s);
        if (z != g_zero.end()) {
            (void)hipFree(z->second);
            g_zero.erase(z);
        }
    }
    HYD_TRY_HIP(hipStreamDestroy(s));
    return HYD_OK;
}

int hydhip_stream_sync(void* stream) {
    HYD_TRY(ensure_init());
    HYD_TRY_HIP(hipStreamSynchronize(resolve_stream(stream)));
    return HYD_OK;
}

int hydhip_device_sync(void) {
    HYD_TRY(ensure_init());
    HYD_TRY_HIP(hipDeviceSynchronize());
    return HYD_OK;
}

void* hydhip_default_stream(void) {
    if (ensure_init() != HYD_OK) return nullptr;
    return g_stream;
}

void* hydhip_scratch(int slot, int64_t bytes, void* stream) {
    if (ensure_init() != HYD_OK) return nullptr;
    if (slot < 0 || slot >= 4 || bytes < 0) {
        set_error(HYD_E_INVALID, "hydhip_scratch: slot %d, %lld bytes", slot, (long long)bytes);
        return nullptr;
    }
    void* p = stream_slot(resolve_stream(stream), slot, (size_t)bytes);
    if (!p) set_error(HYD_E_OOM, "hydhip_scratch: %lld bytes failed", (long long)bytes);
    return p;
}

void* hydhip_zero_scratch(void* stream) {
    if (ensure_init() != HYD_OK) return nullptr;
    void* p = stream_zero_scratch(resolve_stream(stream));
    if (!p) set_error(HYD_E_OOM, "hydhip_zero_scratch: allocation failed");
    return p;
}

int hydhip_event_create(void** event) {
    HYD_TRY(ensure_init());
    hipEvent_t e = nullptr;
    HYD_TRY_HIP(hipEventCreate(&e));
    *event = e;
    return HYD_OK;
}

int hydhip_event_destroy(void* event) {
    if (event) HYD_TRY_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return HYD_OK;
}

int hydhip_event_record(void* event, void* stream) {
    HYD_TRY_HIP(hipEventRecord(static_cast<hipEvent_t>(event), resolve_stream(stream)));
    return HYD_OK;
}

int hydhip_event_elapsed_ms(void* start, void* stop, float* ms) {
    HYD_TRY_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    HYD_TRY_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                    static_cast<hipEvent_t>(stop)));
    return HYD_OK;
}

int hydhip_fill_splitmix_i64(int64_t* dst, int64_t n, uint64_t seed, uint64_t counter0, int shift,
                             void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || shift < 0 || shift > 63 || (n > 0 && !dst))
        return set_error(HYD_E_INVALID, "hydhip_fill_splitmix_i64: bad arguments");
    if (n == 0) return HYD_OK;
    int64_t blocks = ceil_div(n, 256);
    if (blocks > 8192) blocks = 8192;
    fill_splitmix_i64_kernel<<<(unsigned)blocks, 256, 0, resolve_stream(stream)>>>(dst, n, seed,
                                                                                  counter0, shift);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

int hydhip_fill_splitmix_f64(double* dst, int64_t n, uint64_t seed, uint64_t counter0, double scale,
                             double offset, void* stream) {
    HYD_TRY(ensure_init());
    if (n < 0 || (n > 0 && !dst))
        return set_error(HYD_E_INVALID, "hydhip_fill_splitmix_f64: bad arguments");
    if (n == 0) return HYD_OK;
    int64_t blocks = ceil_div(n, 256);
    if (blocks > 8192) blocks = 8192;
    fill_splitmix_f64_kernel<<<(unsigned)blocks, 256, 0, resolve_stream(stream)>>>(
        dst, n, seed, counter0, scale, offset);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

}  // extern "C"
