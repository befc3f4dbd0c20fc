/* hydrangea_runtime.h - the array ABI of the MI355X backend.
 *
 * Drop-in for the reference header runtime/hydrangea_runtime.h: the types
 * hyd_tuple_t, hyd_array_storage_t and hyd_array_t keep the reference layout
 * byte for byte (runtime/hydrangea_runtime.h:220-235), and every function
 * below keeps its name, argument meaning and error behaviour
 * (runtime/hydrangea_runtime.h:237-456).  The difference is where the
 * functions live: the reference defines them `static` in the header, here
 * they are exported by libhydhip.so, which also attaches an HBM mirror to each
 * storage block (see hyd_hip_runtime.h).  A host that calls hyd_export_*
 * kernels therefore includes this header instead of the reference one and
 * links -lhydhip.
 *
 * Errors follow the reference: a message on stderr and exit(1).
 */
#ifndef HYDRANGEA_RUNTIME_H
#define HYDRANGEA_RUNTIME_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
#include <atomic>
#define HYD_ATOMIC_INT64 std::atomic<int64_t>
extern "C" {
#else
#define HYD_ATOMIC_INT64 _Atomic int64_t
#endif

#define HYD_MAX_DIMS 8

/* runtime/hydrangea_runtime.h:220-235 */
typedef struct {
    int64_t elems[HYD_MAX_DIMS];
    int ndims;
} hyd_tuple_t;

typedef struct {
    unsigned char* data;
    HYD_ATOMIC_INT64 refcount;
} hyd_array_storage_t;

typedef struct {
    hyd_array_storage_t* storage;
    unsigned char* data;
    size_t elem_size;
    hyd_tuple_t shape;
} hyd_array_t;

/* Shapes: runtime/hydrangea_runtime.h:237-290 */
hyd_tuple_t hyd_tuple_make(int n, ...);
int64_t hyd_shape_size(hyd_tuple_t shape);
hyd_tuple_t hyd_shape_init(hyd_tuple_t shape);
int64_t hyd_shape_last(hyd_tuple_t shape);
hyd_tuple_t hyd_flat_to_nd(int64_t flat, hyd_tuple_t shape);
int64_t hyd_nd_to_flat(hyd_tuple_t nd, hyd_tuple_t shape);

/* Arrays: runtime/hydrangea_runtime.h:292-368.  Storage is zero-filled
 * (calloc semantics); views share storage and bump the refcount; the last
 * hyd_array_free releases host memory and the HBM mirror. */
hyd_array_t* hyd_array_alloc_bytes(hyd_tuple_t shape, size_t elem_size);
hyd_array_t* hyd_array_alloc_copy(const hyd_array_t* src);
hyd_array_t* hyd_array_alloc(hyd_tuple_t shape);
hyd_array_t* hyd_array_reshape_view(hyd_array_t* arr, hyd_tuple_t shape);
void hyd_array_free(hyd_array_t* arr);
int64_t hyd_array_length(hyd_array_t* arr);
int64_t hyd_array_ndims(hyd_array_t* arr);
int64_t hyd_array_dim(hyd_array_t* arr, int dim);
int64_t* hyd_array_data_int64(hyd_array_t* arr);
double* hyd_array_data_f64(hyd_array_t* arr);

/* sort_indices: runtime/hydrangea_runtime.h:428-456.  Stable ascending sort
 * of a 1-D int64 key array, returning the permutation.  Runs on the GPU. */
hyd_array_t* hyd_sort_indices(hyd_array_t* keys);

/* CSV files: runtime/hydrangea_runtime.h:549-594 (read_array_float,
 * read_array) and runtime/hyd_write_csv.c:24-59 (write_array,
 * write_array_float).  Same names, arguments, text and messages; a missing
 * file or a text the reference rejects prints the reference's message and
 * exits with status 1.  The text is parsed / formatted on the GPU; read
 * results are device-resident (and copied to arr->data under host sync). */
hyd_array_t* hyd_read_array_csv(const char* filename, hyd_tuple_t shape);
hyd_array_t* hyd_read_float_array_csv(const char* filename, hyd_tuple_t shape);
int64_t hyd_write_array_csv(hyd_array_t* arr, const char* filename);
int64_t hyd_write_array_csv_float(hyd_array_t* arr, const char* filename);

/* hyd_get_env_int: runtime/hydrangea_runtime.h:601-615 (unset or
 * unparsable variable: message on stderr, exit(1)). */
int64_t hyd_get_env_int(const char* name);
/* hyd_get_env_string: runtime/hydrangea_runtime.h:617-624. */
const char* hyd_get_env_string(const char* name);
/* getenv_int / getenv_string: runtime/hyd_write_csv.c:61-72 (called by the
 * emitted C for get_env_int / get_env_string, Lowering.hs:1921-1930; atoll;
 * unset variable: message on stderr, exit(1)). */
int64_t getenv_int(const char* name);
const char* getenv_string(const char* name);

/* Slices, element access and printing: runtime/hydrangea_runtime.h:277-290,
 * 458-548 (same output text).  Readers of arr->data copy a device-resident
 * array's bytes back first. */
hyd_tuple_t hyd_slice_shape(hyd_tuple_t src_shape, hyd_tuple_t slice_specs);
double hyd_array_get_float(hyd_array_t* arr, int64_t idx);
void hyd_array_set_float(hyd_array_t* arr, int64_t idx, double v);
void hyd_print_int_inline(int64_t v);
void hyd_print_int(int64_t v);
void hyd_print_double_inline(double v);
void hyd_print_double(double v);
void hyd_print_array_inline(hyd_array_t* arr);
void hyd_print_array(hyd_array_t* arr);
void hyd_print_float_array_inline(hyd_array_t* arr);
void hyd_print_float_array(hyd_array_t* arr);
void hyd_print_tuple(hyd_tuple_t t);
#ifndef __cplusplus
/* Pair printing (C11 _Generic), as the reference header defines it. */
#define _hyd_print_val(x) _Generic((x), \
    double: printf("%.17g", (double)(x)), \
    float: printf("%.17g", (double)(x)), \
    default: printf("%lld", (long long)(x)))
#define hyd_print_pair(a, b) do { \
    printf("("); _hyd_print_val(a); printf(", "); _hyd_print_val(b); printf(")\n"); \
} while (0)
#endif

#ifdef __cplusplus
}
#endif

#endif /* HYDRANGEA_RUNTIME_H */
