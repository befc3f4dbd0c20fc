/* hyd_hip.h - kernel-level C ABI of libhydhip.so (MI355X / gfx950).
 *
 * These are the fused-kernel templates that the HIP backend instantiates for
 * the loop shapes the reference's C backend emits.  Every entry point takes
 * plain device pointers (HBM addresses) and element counts, enqueues its work
 * on `stream` (an opaque hipStream_t; NULL selects the library's own stream)
 * and returns a status code instead of exiting, so a host that cannot afford
 * exit(1) (the serving loop, the test harness) can recover.  The array-level
 * drop-in wrappers (hyd_export_*) are declared in hyd_exports.h.
 *
 * Integer arithmetic is two's-complement mod 2^64 (the reference's C with
 * -fwrapv); fp64 elementwise and stencil arithmetic is IEEE binary64 with no
 * contraction, so results equal the emitted C bit for bit; fp64 reductions
 * use a fixed-shape tree (deterministic run to run).
 */
#ifndef HYD_HIP_H
#define HYD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define HYD_OK 0
#define HYD_E_INVALID 1   /* bad argument: null pointer, negative size, ... */
#define HYD_E_OOM 2       /* device allocation failed */
#define HYD_E_HIP 3       /* HIP runtime error */
#define HYD_E_NODEV 4     /* no GPU */
#define HYD_E_REJECTED 5  /* serving: request can never fit the HBM budget */
#define HYD_E_TIMEOUT 6   /* serving: deadline passed */
#define HYD_E_BUSY 7      /* serving: queue full */

/* Message of the last failing call on this thread (empty if none). */
const char* hydhip_last_error(void);

/* ---- lifecycle: the hyd_metal_init/hyd_metal_cleanup analog of
 *      src/Language/Hydrangea/CodegenMSL.hs:1219-1221 ------------------- */
int hyd_hip_init(const char* code_object_dir_or_null);
void hyd_hip_cleanup(void);
/* Compute units, HBM bytes, LDS bytes per workgroup and device name. */
int hydhip_device_info(int* num_cus, int64_t* hbm_bytes, int64_t* lds_bytes,
                       char* name, int name_len);

/* ---- memory, streams, events (plain pointers) ------------------------- */
int hydhip_malloc(void** dptr, int64_t bytes);
int hydhip_free(void* dptr);
int hydhip_memcpy_h2d(void* dst, const void* src, int64_t bytes, void* stream);
int hydhip_memcpy_d2h(void* dst, const void* src, int64_t bytes, void* stream);
int hydhip_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream);
int hydhip_stream_create(void** stream);
int hydhip_stream_destroy(void* stream);
int hydhip_stream_sync(void* stream);
int hydhip_device_sync(void);
void* hydhip_default_stream(void);
/* Device buffers for generated kernels (hydgen launchers): a buffer of at
 * least `bytes` per (stream, slot), slot 0-3, reused by later calls on the
 * same stream (kernels on one stream run in order).  NULL on failure. */
void* hydhip_scratch(int slot, int64_t bytes, void* stream);
/* A 1 MiB device block per stream that is all zero between kernels: a kernel
 * that adds into it leaves it zero again before it finishes.  NULL on failure. */
void* hydhip_zero_scratch(void* stream);
int hydhip_event_create(void** event);
int hydhip_event_destroy(void* event);
int hydhip_event_record(void* event, void* stream);
int hydhip_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- synthetic inputs: SplitMix64 (Vigna), counter form ---------------
 * word(i) = mix(seed + (counter0 + i + 1) * 0x9E3779B97F4A7C15), which is the
 * (counter0+i)-th output of the sequential generator seeded with `seed`. */
/* dst[i] = word(i) >> shift  (uniform integers in [0, 2^(64-shift))) */
int hydhip_fill_splitmix_i64(int64_t* dst, int64_t n, uint64_t seed,
                             uint64_t counter0, int shift, void* stream);
/* dst[i] = (word(i) >> 11) * 2^-53 * scale + offset, each step exact for
 * scale in {1, 2} and offset in {0, -1}: uniform fp64 in [0,1) or [-1,1). */
int hydhip_fill_splitmix_f64(double* dst, int64_t n, uint64_t seed,
                             uint64_t counter0, double scale, double offset,
                             void* stream);

/* ---- scatter (+) ------------------------------------------------------
 * Strategy: the privatization choice the reference makes in
 * src/Language/Hydrangea/Parallelize.hs:466-503 (privatized) and
 * :350-419 (atomic), remapped to CDNA4 memory levels. */
#define HYD_SCATTER_AUTO 0        /* autotuner / size rule */
#define HYD_SCATTER_LDS 1         /* LDS-private bins per workgroup, one merge */
#define HYD_SCATTER_GLOBAL 2      /* global 64-bit atomics per element */
#define HYD_SCATTER_LDS_SPLIT 3   /* bins split into LDS-sized ranges, one pass per range */

/* Weighted histogram, the fused scatter_generate of
 * `scatter (+) (fill [nbins] 0) (map (x % nbins) data) (map (x * x) data)`:
 * hist[data[i] % nbins] += data[i] * data[i] (mod 2^64), hist zeroed first.
 * Elements whose route falls outside [0, nbins) are dropped. */
int hydhip_weighted_histogram_i64(const int64_t* data, int64_t n, int64_t nbins,
                                  int64_t* hist, int strategy, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc): hydhip_weighted_histogram_i64 with the same arguments
 * and results, whose LDS form merges through a fixed 64 KiB device array
 * (8192 words, one per GPU, held until the process ends) instead of an
 * allocated block.  It calls no hipMalloc/hipFree or allocating routine; the
 * library must be initialised first (hyd_hip_init), else HYD_E_INVALID.
 * HYD_SCATTER_LDS with nbins > 8192 is rejected (HYD_E_INVALID); the automatic
 * choice then takes a form without scratch.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (other streams, other threads) race. */
int hydhip_weighted_histogram_i64_noalloc(const int64_t* data, int64_t n, int64_t nbins,
                                          int64_t* hist, int strategy, void* stream);
/* Guarded scatter (+) on int64, scatter_guarded of
 * src/Language/Hydrangea/Lowering.hs:1432-1517: dst holds the defaults and
 * receives dst[idx[i]] += vals[i] for every i with guard[i] != 0 (guard NULL:
 * every i).  Routes outside [0, ndst) are dropped. */
int hydhip_scatter_guarded_add_i64(const int64_t* idx, const int64_t* vals,
                                   const int64_t* guard, int64_t n, int64_t* dst,
                                   int64_t ndst, int strategy, void* stream);
/* Scatter chain (+) on int64 (EScatterChain, Lowering.hs:1622-1858): phase k
 * is a guarded scatter of lens[k] elements into the same dst, in turn.
 * idx/vals/guard/lens are HOST arrays of device pointers and lengths;
 * guard (or guard[k]) NULL means unguarded. */
int hydhip_scatter_chain_add_i64(int64_t nphases, const int64_t* const* idx,
                                 const int64_t* const* vals, const int64_t* const* guard,
                                 const int64_t* lens, int64_t* dst, int64_t ndst,
                                 int strategy, void* stream);
/* The same on fp64 ((+.)), bit-exact to the reference's serial loops: per
 * bin, the values arrive in element order (chains: phase by phase), which a
 * stable sort by bin reproduces before an in-order fold per bin. */
int hydhip_scatter_guarded_add_f64(const int64_t* idx, const double* vals,
                                   const int64_t* guard, int64_t n, double* dst,
                                   int64_t ndst, void* stream);
int hydhip_scatter_chain_add_f64(int64_t nphases, const int64_t* const* idx,
                                 const double* const* vals, const int64_t* const* guard,
                                 const int64_t* lens, double* dst, int64_t ndst,
                                 void* stream);
/* Unfused scatter: dst[idx[i]] += vals[i] (dst is accumulated, not zeroed). */
int hydhip_scatter_add_i64(const int64_t* idx, const int64_t* vals, int64_t n,
                           int64_t* dst, int64_t ndst, int strategy, void* stream);

/* ---- reductions and maps (fp64) ---------------------------------------- */
/* *out = init + sum_i xs[i]*ys[i]  (the fused reduce_generate of
 * `reduce (+.) init (zipwith (*.) xs ys)`), fixed-shape tree. out is device. */
int hydhip_dot_f64(const double* xs, const double* ys, int64_t n, double init,
                   double* out, void* stream);
/* The fused 8-op chain, out[i] = f(xs[i], ys[i]) with
 *   t1 = 2.0*x; t2 = t1+y; t3 = |t2|; t4 = sqrt t3; t5 = t4*x; t6 = t5-y;
 *   t7 = t6 > 0.0 ? t6 : 0.0; out = exp(t7*1.0e-3). */
int hydhip_map_chain_f64(const double* xs, const double* ys, int64_t n,
                         double* out, void* stream);

/* The map chain's math on its own: out[i] = exp(x[i]) (op 0, the libm-exact
 * table method of the chain) or sqrt(x[i]) (op 1), on the GPU. */
int hydhip_math_f64(int op, const double* x, int64_t n, double* out, void* stream);
/* The same exp on the host CPU (identical operation sequence). */
void hydhip_host_exp(const double* x, int64_t n, double* out);
/* The histograms' route x % m (C truncating remainder, m >= 1) by the
 * division-free method the kernels use, run on the host CPU: out[i] = x[i] % m. */
int hydhip_host_mod_i64(const int64_t* x, int64_t n, int64_t m, int64_t* out);
/* Streaming copy dst[i] = src[i] (n even, 16-byte aligned): the bandwidth
 * reference for the fused kernels. */
int hydhip_copy_f64(const double* src, int64_t n, double* dst, void* stream);

/* ---- 2-D clamp stencils (fp64, row-major h x w) ------------------------ */
#define HYD_STENCIL_HEAT 0       /* c + alpha*((((n + s) + w) + e) - 4.0*c) */
#define HYD_STENCIL_LAPLACIAN 1  /* (((n + s) + w) + e) + c*(-4.0) */
#define HYD_STENCIL_JACOBI4 2    /* (((n + s) + w) + e) / 4.0 */
#define HYD_STENCIL_AVG5 3       /* ((((n + s) + w) + e) + c) / 5.0 */
/* n = a(-1,0), s = a(1,0), w = a(0,-1), e = a(0,1), c = a(0,0), indices
 * clamped to the grid.  alpha is used by HYD_STENCIL_HEAT only. */
int hydhip_stencil2d_clamp_f64(const double* in, int64_t h, int64_t w, int body,
                               double alpha, double* out, void* stream);
/* iterate iters (stencil clamp body): ping-pong between buf0 and buf1, init is
 * never written; *result receives buf0 or buf1 (buf0 when iters == 0, after
 * a copy).  use_graph != 0 captures the iters launches once as a hipGraph and
 * replays it (cached per pointer/shape/iters). */
int hydhip_stencil2d_iterate_f64(const double* init, int64_t h, int64_t w,
                                 int body, double alpha, int64_t iters,
                                 double* buf0, double* buf1, int use_graph,
                                 double** result, void* stream);

/* ---- batched scatter ------------------------------------------------------
 * Many weighted-histogram instances in one persistent launch.  Instance k
 * reads data[off[k] .. off[k]+len[k]) and writes hist[hoff[k] ..
 * hoff[k]+bins[k]) (zeroed first): hist[x % bins] += x*x.  The descriptor
 * arrays off/len/bins/hoff are HOST memory (the plan is built on the host).
 * policy: 0 = autotuned per size class, 1 = LDS bins wherever they fit,
 * 2 = global atomics everywhere. */
int hydhip_batched_histogram_i64(const int64_t* data, const int64_t* off,
                                 const int64_t* len, const int64_t* bins,
                                 const int64_t* hoff, int64_t count, int64_t* hist,
                                 int policy, void* stream);
/* Autotuner of the batched scatter: measures LDS vs global atomics for each
 * (bin class, size class) and stores the table at `path` (NULL: the
 * HYD_TUNE_CACHE variable or ~/.cache/hydhip/tune_<arch>.txt).  Returns the
 * number of classes measured, or a negative status. */
int hydhip_tune_batched(const char* path, void* stream);
int hydhip_tune_set(int64_t bins, int64_t n, int mode);
int hydhip_tune_clear(void);
int hydhip_tune_table(int* out, int max_entries);
/* Launch shapes: the streaming grid (workgroups per CU) of the map chain and
 * the copy, per size class (floor log2 n); a call takes the largest tuned
 * class <= its own, else the flat form (HYD_SHAPE_FLAT).  hydhip_tune_shapes
 * measures the flat form and 1, 2, 4 and 8 per CU at 2^20, 2^23 and 2^26
 * elements and stores the choices
 * in the same table file as hydhip_tune_batched; hydhip_tune_shape_set sets
 * one class (per_cu <= 0 removes it); hydhip_tune_shape_table copies
 * (kind, n_log2, per_cu) triples.  The dot keeps a fixed grid: its partial
 * count fixes the order of its sum.  No reference counterpart: the
 * reference's loops have no launch shape. */
#define HYD_SHAPE_MAP_CHAIN 0
#define HYD_SHAPE_COPY 1
#define HYD_SHAPE_KINDS 2
/* per_cu value of the flat form: no loop, one contiguous block of pairs per
 * workgroup, the grid sized by the data (the default of untuned classes). */
#define HYD_SHAPE_FLAT 65536
int hydhip_tune_shapes(const char* path, void* stream);
int hydhip_tune_shape_set(int kind, int64_t n, int per_cu);
int hydhip_tune_shape_table(int* out, int max_entries);

/* Many dot products in one launch: *outs[r] = init + sum_i xs[r][i] * ys[r][i]
 * for requests r < count (the zipwith -> reduce of programs/dot.hyd, one
 * instance per request).  xs, ys, lens and outs are HOST arrays of device
 * pointers and lengths.  Each request is cut into fixed 65536-element tiles
 * summed in a fixed order, so its bits do not depend on the other requests
 * of the batch (results within 1e-12 relative of the serial sum). */
int hydhip_batched_dot_f64(const double* const* xs, const double* const* ys,
                           const int64_t* lens, int64_t count, double init,
                           double* const* outs, void* stream);

/* ---- scans, sort, gather, COO -> CSR (int64) ---------------------------- */
/* The scan family of src/Language/Hydrangea/Lowering.hs:3563-3607 with (+):
 * exclusive: out[k] = init + sum_{j<k} a[j]; inclusive adds a[k] too;
 * reverse runs from the end (scanr).  out may alias in. */
int hydhip_scan_i64(const int64_t* in, int64_t n, int64_t init, int64_t* out,
                    int inclusive, int reverse, void* stream);
/* sort_indices (runtime/hydrangea_runtime.h:391-456): stable LSD radix sort
 * of keys, 8-bit digits, passes = bytes of max(0, max key) (at least one);
 * perm[k] = index of the k-th smallest key.  Reads the maximum key back to
 * the host (one synchronisation). */
int hydhip_sort_indices_i64(const int64_t* keys, int64_t n, int64_t* perm,
                            void* stream);
/* gather (Lowering.hs:1860-1898): out[i] = src[perm[i]]; an index outside
 * [0, nsrc) yields 0. */
int hydhip_gather_i64(const int64_t* perm, int64_t n, const int64_t* src,
                      int64_t nsrc, int64_t* out, void* stream);
/* segmented_reduce (+) (Lowering.hs:874-938): out[s] = init + sum of
 * vals[offsets[s] .. offsets[s+1]), s < noffsets-1. */
int hydhip_segmented_reduce_i64(const int64_t* offsets, int64_t noffsets,
                                const int64_t* vals, int64_t nvals, int64_t init,
                                int64_t* out, void* stream);
/* coo_sum_duplicates (Lowering.hs:966-1067): merges adjacent equal
 * (row, col) of the first nnz triplets; outputs are len long, zero tail;
 * *nnz_out (host) receives the merged count. */
int hydhip_coo_sum_duplicates_i64(const int64_t* rows, const int64_t* cols,
                                  const int64_t* vals, int64_t nnz, int64_t len,
                                  int64_t* out_rows, int64_t* out_cols,
                                  int64_t* out_vals, int64_t* nnz_out,
                                  void* stream);
/* csr_from_sorted_coo (Lowering.hs:1068-1147): row_ptr[nrows+1] (exclusive
 * scan of per-row counts of the first nnz rows, row_ptr[nrows] = total),
 * col_idx/out_vals = first nnz of cols/vals then zeros up to len.  Rows
 * outside [0, nrows) are not counted. */
int hydhip_csr_from_sorted_coo_i64(int64_t nrows, int64_t nnz, const int64_t* rows,
                                   const int64_t* cols, const int64_t* vals,
                                   int64_t len, int64_t* row_ptr, int64_t* col_idx,
                                   int64_t* out_vals, void* stream);
/* The COO -> CSR build program (bench/coo_csr_build/coo_csr_build.hyd:44-62):
 * keys = row*ncols + col, perm = sort_indices keys, gather, merge duplicates,
 * CSR.  row_ptr[nrows+1], col_idx[nnz] and out_vals[nnz] (zero past
 * *nnz_out); *nnz_out (host) = merged count. */
int hydhip_coo_to_csr_i64(int64_t nrows, int64_t ncols, int64_t nnz,
                          const int64_t* rows, const int64_t* cols,
                          const int64_t* vals, int64_t* row_ptr, int64_t* col_idx,
                          int64_t* out_vals, int64_t* nnz_out, void* stream);

/* ---- CSV text <-> arrays (device buffers) ------------------------------ */
/* The element reads of read_array / read_array_float
 * (runtime/hydrangea_runtime.h:549-594): text[0, nbytes) in device memory
 * holds numbers as fscanf "%lld" / "%lf" reads them (glibc rules, correctly
 * rounded fp64), separated as the reference's loop skips them (',' ' ' '\n'
 * '\r', then isspace).  The first `count` values go to out (device).  A text
 * the reference would reject is reported in *status (host), not by exit:
 * error_kind HYDHIP_CSV_PARSE_ERROR ("parse error at element N") or
 * HYDHIP_CSV_TOO_FEW ("too few values"), at error_element.  Synchronises the
 * stream (the status is read back). */
#define HYDHIP_CSV_OK 0
#define HYDHIP_CSV_PARSE_ERROR 1
#define HYDHIP_CSV_TOO_FEW 2
typedef struct {
    int64_t error_element; /* first element the reference fails on, or -1 */
    int32_t error_kind;    /* HYDHIP_CSV_OK / _PARSE_ERROR / _TOO_FEW */
    int32_t reserved;
    int64_t numbers_found; /* numbers read before the first failure */
    int64_t exact_path;    /* fp64 numbers rounded by the big-integer path */
} hydhip_csv_status_t;
int hydhip_csv_parse_i64(const void* text, int64_t nbytes, int64_t count,
                         int64_t* out, hydhip_csv_status_t* status, void* stream);
int hydhip_csv_parse_f64(const void* text, int64_t nbytes, int64_t count,
                         double* out, hydhip_csv_status_t* status, void* stream);
/* The row writers of write_array / write_array_float
 * (runtime/hyd_write_csv.c:5-59): fprintf "%lld" / "%.17g" text of `count`
 * values, ',' between the elements of a row and '\n' after every row_len-th
 * (row_len = columns of a 2-D array, else count), into text (device,
 * `capacity` bytes; hydhip_csv_text_bound(count) always suffices).  *nbytes
 * (host) receives the length; *exact_path (host, may be NULL) the fp64
 * values rounded by the big-integer path.  Synchronises the stream. */
int64_t hydhip_csv_text_bound(int64_t count);
int hydhip_csv_format_i64(const int64_t* vals, int64_t count, int64_t row_len,
                          void* text, int64_t capacity, int64_t* nbytes,
                          void* stream);
int hydhip_csv_format_f64(const double* vals, int64_t count, int64_t row_len,
                          void* text, int64_t capacity, int64_t* nbytes,
                          int64_t* exact_path, void* stream);
/* CPU test hooks: the same conversions compiled for the host (no GPU), for
 * checking the algorithms against the C library at volumes the GPU tests do
 * not reach.  Not part of the product path.
 * scan: fscanf-like on s[0, len): skips isspace, reads one number; returns 1
 * (value, *consumed = characters used), 0 (no number), -1 (only space).
 * format: text of each value at out + k * stride (no separator), lengths in
 * lens; returns how many fp64 values took the big-integer path.
 * flags & 1 sends every fp64 value through the big-integer path. */
/* Test hook: on = 1 formats every fp64 value through the exact (big-integer)
 * kernel, which real values almost never need, so that its GPU path can be
 * checked; on = 0 restores the normal path.  Not for production use. */
void hydhip_csv_test_force_exact(int on);
/* CPU test hook: out[i] = 1 when bytes[i] is a gap character (',' or
 * isspace) by the 8-byte classifier of the parse kernels; n % 8 == 0. */
void hydhip_host_csv_gap_flags(const void* bytes, int64_t n, unsigned char* out);
int hydhip_host_csv_scan(const char* s, int64_t len, int is_float, int flags,
                         void* value, int64_t* consumed, int* exact);
int64_t hydhip_host_csv_format(const void* vals, int64_t n, int is_float,
                               int flags, char* out, int64_t stride,
                               int32_t* lens);

#ifdef __cplusplus
}
#endif

#endif /* HYD_HIP_H */
