// arrays.hip - the hyd_array_t runtime (runtime/hydrangea_runtime.h:220-456
// of the reference) with an HBM mirror per storage block.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <unordered_set>

#include "../../include/hyd_hip_runtime.h"
#include "hyd_internal.hpp"

extern "C" void hydhip_forget_graphs_for(const void* ptr);

namespace hyd {
namespace {

constexpr uint64_t kMagic = 0x4859444849505354ULL;
// host <-> HBM copies of at least this size go through the pinned halves
constexpr size_t kStagedBytes = (size_t)8 << 20;

// The storage block handed out through hyd_array_t::storage.  The reference
// layout {data, refcount} comes first, so code that only knows the reference
// struct reads it unchanged.
struct HipStorage {
    hyd_array_storage_t base;
    uint64_t magic = kMagic;
    size_t bytes = 0;
    void* dev = nullptr;           // HBM copy, allocated on first use
    bool device_resident = false;  // HBM copy is authoritative
    bool host_valid = true;        // host bytes match the HBM copy
};

std::mutex g_mu;
std::unordered_set<hyd_array_storage_t*> g_storages;
int g_host_sync = 1;
void* g_export_stream = nullptr;

[[noreturn]] void die(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "hydrangea: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    exit(1);
}

HipStorage* storage_of(hyd_array_t* arr) {
    if (!arr || !arr->storage) die("null array");
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_storages.count(arr->storage)) die("array storage %p was not allocated by libhydhip", (void*)arr->storage);
    }
    return reinterpret_cast<HipStorage*>(arr->storage);
}

hyd_array_t* make_header(HipStorage* st, hyd_tuple_t shape, size_t elem_size) {
    hyd_array_t* arr = (hyd_array_t*)malloc(sizeof(hyd_array_t));
    if (!arr) die("header allocation failed");
    arr->storage = &st->base;
    arr->data = st->base.data;
    arr->elem_size = elem_size;
    arr->shape = shape;
    return arr;
}

HipStorage* new_storage(size_t bytes) {
    HipStorage* st = new HipStorage();
    // calloc: zero-filled like the reference (hydrangea_runtime.h:304); large
    // blocks come from fresh zero pages, so an unused host copy costs nothing.
    st->base.data = (unsigned char*)calloc(bytes ? bytes : 1, 1);
    if (!st->base.data) die("host allocation of %zu bytes failed", bytes);
    st->base.refcount = 1;
    st->bytes = bytes;
    std::lock_guard<std::mutex> lk(g_mu);
    g_storages.insert(&st->base);
    return st;
}

int ensure_dev(HipStorage* st) {
    if (st->dev || st->bytes == 0) return HYD_OK;
    HYD_TRY(ensure_init());
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, st->bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(HYD_E_OOM, "hipMalloc(%zu) failed: %s", st->bytes, hipGetErrorString(e));
    }
    st->dev = p;
    return HYD_OK;
}

}  // namespace

// Shared with exports.hip.
hipStream_t export_stream() { return resolve_stream(g_export_stream); }
int host_sync_enabled() { return g_host_sync; }

}  // namespace hyd

using namespace hyd;

namespace hyd {
void mark_device_written(hyd_array_t* arr);
}

extern "C" {

hyd_tuple_t hyd_tuple_make(int n, ...) {
    hyd_tuple_t t;
    memset(&t, 0, sizeof t);
    t.ndims = n;
    va_list args;
    va_start(args, n);
    for (int i = 0; i < n && i < HYD_MAX_DIMS; i++) t.elems[i] = va_arg(args, int64_t);
    va_end(args);
    return t;
}

int64_t hyd_shape_size(hyd_tuple_t shape) {
    int64_t size = 1;
    for (int i = 0; i < shape.ndims; i++) size *= shape.elems[i];
    return size;
}

hyd_tuple_t hyd_shape_init(hyd_tuple_t shape) {
    hyd_tuple_t out;
    memset(&out, 0, sizeof out);
    if (shape.ndims <= 0) return out;
    out.ndims = shape.ndims - 1;
    for (int i = 0; i < out.ndims; i++) out.elems[i] = shape.elems[i];
    return out;
}

int64_t hyd_shape_last(hyd_tuple_t shape) {
    return shape.ndims <= 0 ? 0 : shape.elems[shape.ndims - 1];
}

hyd_tuple_t hyd_flat_to_nd(int64_t flat, hyd_tuple_t shape) {
    hyd_tuple_t idx;
    memset(&idx, 0, sizeof idx);
    idx.ndims = shape.ndims;
    int64_t rem = flat;
    for (int i = shape.ndims - 1; i >= 0; i--) {
        idx.elems[i] = rem % shape.elems[i];
        rem /= shape.elems[i];
    }
    return idx;
}

int64_t hyd_nd_to_flat(hyd_tuple_t nd, hyd_tuple_t shape) {
    int64_t flat = 0, stride = 1;
    for (int i = shape.ndims - 1; i >= 0; i--) {
        flat += nd.elems[i] * stride;
        stride *= shape.elems[i];
    }
    return flat;
}

hyd_array_t* hyd_array_alloc_bytes(hyd_tuple_t shape, size_t elem_size) {
    int64_t size = hyd_shape_size(shape);
    if (size < 0) die("negative array size %lld", (long long)size);
    HipStorage* st = new_storage((size_t)size * elem_size);
    return make_header(st, shape, elem_size);
}

hyd_array_t* hyd_array_alloc(hyd_tuple_t shape) { return hyd_array_alloc_bytes(shape, sizeof(int64_t)); }

hyd_array_t* hyd_array_alloc_copy(const hyd_array_t* src) {
    hyd_array_t* s = const_cast<hyd_array_t*>(src);
    HipStorage* ss = storage_of(s);
    hyd_array_t* dst = hyd_array_alloc_bytes(src->shape, src->elem_size);
    size_t bytes = (size_t)hyd_shape_size(src->shape) * src->elem_size;
    if (ss->device_resident && !ss->host_valid) {
        // copy HBM to HBM; the new array is device-resident as well
        HipStorage* ds = storage_of(dst);
        if (ensure_dev(ds) != HYD_OK) die("%s", hydhip_last_error());
        size_t off = (size_t)(src->data - ss->base.data);
        hipStream_t st = export_stream();
        if (bytes && hipMemcpyAsync(ds->dev, (unsigned char*)ss->dev + off, bytes,
                                    hipMemcpyDeviceToDevice, st) != hipSuccess)
            die("device copy failed");
        ds->device_resident = true;
        ds->host_valid = false;
        if (g_host_sync) hyd_array_sync_host(dst);
    } else if (bytes) {
        memcpy(dst->data, src->data, bytes);
    }
    return dst;
}

hyd_array_t* hyd_array_reshape_view(hyd_array_t* arr, hyd_tuple_t shape) {
    int64_t src_size = hyd_shape_size(arr->shape);
    int64_t dst_size = hyd_shape_size(shape);
    if (src_size != dst_size) {
        fprintf(stderr, "hydrangea: reshape view size mismatch (%lld elements to %lld)\n",
                (long long)src_size, (long long)dst_size);
        exit(1);
    }
    HipStorage* st = storage_of(arr);
    st->base.refcount += 1;
    hyd_array_t* v = make_header(st, shape, arr->elem_size);
    v->data = arr->data;
    return v;
}

void hyd_array_free(hyd_array_t* arr) {
    if (arr == NULL) return;
    HipStorage* st = storage_of(arr);
    st->base.refcount -= 1;
    if (st->base.refcount == 0) {
        {
            std::lock_guard<std::mutex> lk(g_mu);
            g_storages.erase(&st->base);
        }
        if (st->dev) {
            // drain queued work that may read this buffer (or a graph that
            // names it) before the graph and the memory are released
            (void)hipStreamSynchronize(export_stream());
            hydhip_forget_graphs_for(st->dev);
            (void)hipFree(st->dev);
        }
        free(st->base.data);
        delete st;
    }
    free(arr);
}

int64_t hyd_array_length(hyd_array_t* arr) { return hyd_shape_size(arr->shape); }
int64_t hyd_array_ndims(hyd_array_t* arr) { return arr->shape.ndims; }
int64_t hyd_array_dim(hyd_array_t* arr, int dim) { return arr->shape.elems[dim]; }
int64_t* hyd_array_data_int64(hyd_array_t* arr) { return (int64_t*)(void*)arr->data; }
double* hyd_array_data_f64(hyd_array_t* arr) { return (double*)(void*)arr->data; }

int64_t hyd_get_env_int(const char* name) {
    char* val = getenv(name);
    if (!val) {
        fprintf(stderr, "get_env_int: environment variable '%s' not set\n", name);
        exit(1);
    }
    char* end;
    long result = strtol(val, &end, 10);
    if (*end != '\0' || val == end) {
        fprintf(stderr, "get_env_int: cannot parse '%s' as int\n", val);
        exit(1);
    }
    return result;
}

const char* hyd_get_env_string(const char* name) {
    char* val = getenv(name);
    if (!val) {
        fprintf(stderr, "get_env_string: environment variable '%s' not set\n", name);
        exit(1);
    }
    return val;
}

/* runtime/hyd_write_csv.c:61-72: the helpers the emitted C calls for
 * get_env_int / get_env_string (Lowering.hs:1921-1930).  atoll, as there. */
int64_t getenv_int(const char* name) {
    char* val = getenv(name);
    if (!val) {
        fprintf(stderr, "environment variable %s not set\n", name);
        exit(1);
    }
    return atoll(val);
}

const char* getenv_string(const char* name) {
    char* val = getenv(name);
    if (!val) {
        fprintf(stderr, "environment variable %s not set\n", name);
        exit(1);
    }
    return val;
}

/* ---- slices, element access, printing (runtime/hydrangea_runtime.h:277-290,
 * 458-548): the host helpers emitted programs call.  Readers of arr->data
 * first bring a device-resident array's bytes back. */
hyd_tuple_t hyd_slice_shape(hyd_tuple_t src, hyd_tuple_t specs) {
    hyd_tuple_t out;
    memset(&out, 0, sizeof out);
    out.ndims = src.ndims;
    for (int i = 0; i < src.ndims && i < HYD_MAX_DIMS; i++) {
        const int64_t start = 2 * i < HYD_MAX_DIMS ? specs.elems[2 * i] : -1;
        const int64_t len = 2 * i + 1 < HYD_MAX_DIMS ? specs.elems[2 * i + 1] : -1;
        out.elems[i] = (start == -1 && len == -1) ? src.elems[i] : len;  /* (-1, -1): the whole axis */
    }
    return out;
}

static void host_view(hyd_array_t* arr) {
    if (hyd_array_sync_host(arr) != HYD_OK) die("%s", hydhip_last_error());
}

double hyd_array_get_float(hyd_array_t* arr, int64_t idx) {
    host_view(arr);
    double v;
    memcpy(&v, arr->data + (size_t)idx * arr->elem_size, sizeof v);
    return v;
}

void hyd_array_set_float(hyd_array_t* arr, int64_t idx, double v) {
    host_view(arr);
    memcpy(arr->data + (size_t)idx * arr->elem_size, &v, sizeof v);
    if (hyd_array_is_device_resident(arr)) hyd_array_mark_host_dirty(arr);  /* the host copy is newer now */
}

void hyd_print_int_inline(int64_t v) { printf("%lld", (long long)v); }
void hyd_print_int(int64_t v) { printf("%lld\n", (long long)v); }
void hyd_print_double_inline(double v) { printf("%.17g", v); }
void hyd_print_double(double v) { printf("%.17g\n", v); }

static void print_shape_suffix(const hyd_array_t* arr) {
    printf("] (shape: [");
    for (int i = 0; i < arr->shape.ndims; i++) printf(i ? ", %lld" : "%lld", (long long)arr->shape.elems[i]);
    printf("])");
}

void hyd_print_array_inline(hyd_array_t* arr) {
    host_view(arr);
    const int64_t n = hyd_shape_size(arr->shape);
    printf("[");
    for (int64_t i = 0; i < n; i++) {
        int64_t v;
        memcpy(&v, arr->data + (size_t)i * arr->elem_size, sizeof v);
        printf(i ? ", %lld" : "%lld", (long long)v);
    }
    print_shape_suffix(arr);
}

void hyd_print_array(hyd_array_t* arr) {
    hyd_print_array_inline(arr);
    printf("\n");
}

void hyd_print_float_array_inline(hyd_array_t* arr) {
    host_view(arr);
    const int64_t n = hyd_shape_size(arr->shape);
    printf("[");
    for (int64_t i = 0; i < n; i++) {
        double v;
        memcpy(&v, arr->data + (size_t)i * arr->elem_size, sizeof v);
        printf(i ? ", %.17g" : "%.17g", v);
    }
    print_shape_suffix(arr);
}

void hyd_print_float_array(hyd_array_t* arr) {
    hyd_print_float_array_inline(arr);
    printf("\n");
}

void hyd_print_tuple(hyd_tuple_t t) {
    printf("(");
    for (int i = 0; i < t.ndims; i++) printf(i ? ", %lld" : "%lld", (long long)t.elems[i]);
    printf(")\n");
}

/* ---- residency ---------------------------------------------------------- */

hyd_array_t* hyd_array_alloc_device(hyd_tuple_t shape, size_t elem_size) {
    hyd_array_t* arr = hyd_array_alloc_bytes(shape, elem_size);
    HipStorage* st = storage_of(arr);
    if (ensure_dev(st) != HYD_OK) die("%s", hydhip_last_error());
    if (st->bytes && fill_zero(st->dev, st->bytes, export_stream()) != HYD_OK)
        die("%s", hydhip_last_error());
    st->device_resident = true;
    st->host_valid = true;  // both copies are zero
    return arr;
}

void* hyd_array_device_data(hyd_array_t* arr) {
    HipStorage* st = storage_of(arr);
    if (st->bytes == 0) return nullptr;
    if (ensure_dev(st) != HYD_OK) die("%s", hydhip_last_error());
    if (!st->device_resident) {
        // host copy authoritative: upload (ordered before the kernels that
        // follow on the export stream; large arrays through the pinned halves)
        hipStream_t s = export_stream();
        if (st->bytes >= kStagedBytes) {
            if (staged_h2d(st->dev, st->base.data, st->bytes, s) != HYD_OK)
                die("upload of %zu bytes failed: %s", st->bytes, hydhip_last_error());
        } else if (hipMemcpyAsync(st->dev, st->base.data, st->bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
                   hipStreamSynchronize(s) != hipSuccess) {
            die("upload of %zu bytes failed", st->bytes);
        }
    }
    return (unsigned char*)st->dev + (arr->data - st->base.data);
}

void* hyd_array_device_data_keep(hyd_array_t* arr) {
    HipStorage* st = storage_of(arr);
    void* d = hyd_array_device_data(arr);
    if (!st->device_resident && st->bytes) {  // just uploaded: both copies are current now
        st->device_resident = true;
        st->host_valid = true;
    }
    return d;
}

int hyd_array_enter_host(hyd_array_t* arr) {
    HipStorage* st = storage_of(arr);
    if (st->device_resident && st->host_valid) st->device_resident = false;  // the host bytes may have changed
    return HYD_OK;
}

int hyd_array_sync_host(hyd_array_t* arr) {
    HipStorage* st = storage_of(arr);
    if (!st->device_resident || st->host_valid || st->bytes == 0) return HYD_OK;
    hipStream_t s = export_stream();
    if (st->bytes >= kStagedBytes) {
        HYD_TRY(staged_d2h(st->base.data, st->dev, st->bytes, s));
    } else {
        HYD_TRY_HIP(hipMemcpyAsync(st->base.data, st->dev, st->bytes, hipMemcpyDeviceToHost, s));
        HYD_TRY_HIP(hipStreamSynchronize(s));
    }
    st->host_valid = true;
    return HYD_OK;
}

int hyd_array_mark_host_dirty(hyd_array_t* arr) {
    HipStorage* st = storage_of(arr);
    st->device_resident = false;
    st->host_valid = true;
    return HYD_OK;
}

int hyd_array_is_device_resident(hyd_array_t* arr) { return storage_of(arr)->device_resident ? 1 : 0; }

int hyd_array_mark_device_written(hyd_array_t* arr) {
    hyd::mark_device_written(arr);
    return HYD_OK;
}

void hyd_hip_set_host_sync(int on) { g_host_sync = on ? 1 : 0; }
int hyd_hip_get_host_sync(void) { return g_host_sync; }
void hyd_hip_set_stream(void* stream) { g_export_stream = stream; }
void* hyd_hip_get_stream(void) {
    if (ensure_init() != HYD_OK) return nullptr;
    return export_stream();
}

hyd_array_t* hyd_array_alloc_device8(hyd_tuple_t shape) { return hyd_array_alloc_device(shape, 8); }

}  // extern "C"

namespace hyd {
// A kernel just wrote the HBM copy of `arr`.
void mark_device_written(hyd_array_t* arr) {
    HipStorage* st = storage_of(arr);
    st->device_resident = true;
    st->host_valid = false;
}
}  // namespace hyd
