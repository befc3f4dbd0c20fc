/* hyd_oracle.c - CPU oracle for the fused array kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in C99 + OpenMP, the loop
 * nests the reference's C backend emits for each hot-path kernel, so the GPU
 * results can be checked against them.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py load it; the product (libhydhip.so)
 * never links or calls it.
 *
 * Build (oracle/build.py): gcc -O3 -march=x86-64-v3 -std=c99 -fopenmp
 *   -ffp-contract=off -fwrapv -shared -fPIC
 * which is the reference's `cc -O3 -march=native -std=c99 -fopenmp`
 * (app/CBackend.hs:89) with contraction and signed wraparound made explicit
 * (ISO C99 mode already implies no contraction in GCC).
 *
 * Every function cites the reference code whose semantics it restates.  How
 * each one is pinned (golden vectors, closed forms, brute force) is listed
 * in tests/test_oracle.py; the only function not pinned bit-for-bit is the
 * parallel reduction order of oracle_dot (OpenMP decides it in the
 * reference; it is checked against an exactly rounded sum within 1e-12).
 */
#define _POSIX_C_SOURCE 200809L /* fmemopen, open_memstream */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- SplitMix64 (Vigna), counter form ---------------------------------- */
static inline uint64_t splitmix_at(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void oracle_splitmix_i64(uint64_t seed, uint64_t c0, int64_t n, int shift, int64_t* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = (int64_t)(splitmix_at(seed, c0 + (uint64_t)i) >> shift);
}

void oracle_splitmix_f64(uint64_t seed, uint64_t c0, int64_t n, double scale, double offset,
                         double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double u = (double)(splitmix_at(seed, c0 + (uint64_t)i) >> 11) * 0x1.0p-53;
        double v = u * scale;
        out[i] = v + offset;
    }
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

/* ---- scatter (+) --------------------------------------------------------- */

/* Serial lowering of scatter (+): Lowering.hs:1408-1430 with the combine
 * write of :3553-3557 (old = dst[idx]; new = val + old; dst[idx] = new). */
void oracle_scatter_add_serial(const int64_t* idx, const int64_t* vals, int64_t n, int64_t* dst,
                               int64_t ndst) {
    for (int64_t i = 0; i < n; i++) {
        int64_t b = idx[i];
        if (b < 0 || b >= ndst) continue; /* proven in range by the reference's types */
        int64_t old = dst[b];
        int64_t nw = vals[i] + old;
        dst[b] = nw;
    }
}

/* Privatized scatter (CodegenC.hs:1119-1288): nthreads x ndst private grids
 * (calloc), `omp for` over the elements into the thread's grid, then
 * `omp for` over bins: sum = dst[b]; sum += grid[t][b] for t < nt; dst[b] = sum. */
void oracle_scatter_add_privatized(const int64_t* idx, const int64_t* vals, int64_t n,
                                   int64_t* dst, int64_t ndst) {
    int nthr = omp_get_max_threads();
    int64_t* grids = (int64_t*)calloc((size_t)nthr * (size_t)ndst, sizeof(int64_t));
    if (grids == NULL) {
        fprintf(stderr, "hydrangea: failed to allocate privatized scatter buffer\n");
        exit(1);
    }
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        int nt = omp_get_num_threads();
        int64_t* priv = grids + (size_t)tid * (size_t)ndst;
#pragma omp for
        for (int64_t i = 0; i < n; i++) {
            int64_t b = idx[i];
            if (b < 0 || b >= ndst) continue;
            priv[b] += vals[i];
        }
#pragma omp for
        for (int64_t m = 0; m < ndst; m++) {
            int64_t sum = dst[m];
            for (int t = 0; t < nt; t++) sum += grids[(size_t)t * (size_t)ndst + m];
            dst[m] = sum;
        }
    }
    free(grids);
}

/* Guarded scatter (+), Lowering.hs:1432-1517: dst starts as the defaults
 * (the caller's copy); for i in order: if (guard[i]) { old = dst[idx];
 * new = val + old; dst[idx] = new; }.  A guard is true when nonzero (bools
 * are int64 in the emitted C, CodegenC.hs:2016).  guard == NULL: every
 * element contributes (plain scatter). */
void oracle_scatter_guarded_add_i64(const int64_t* idx, const int64_t* vals, const int64_t* guard,
                                    int64_t n, int64_t* dst, int64_t ndst) {
    for (int64_t i = 0; i < n; i++) {
        if (guard && !guard[i]) continue;
        int64_t b = idx[i];
        if (b < 0 || b >= ndst) continue; /* proven in range by the reference's types */
        int64_t old = dst[b];
        int64_t nw = vals[i] + old;
        dst[b] = nw;
    }
}

/* The same loop over fp64 values ((+.)): order-dependent, so this serial
 * index order is the contract the GPU reproduces bit for bit. */
void oracle_scatter_guarded_add_f64(const int64_t* idx, const double* vals, const int64_t* guard,
                                    int64_t n, double* dst, int64_t ndst) {
    for (int64_t i = 0; i < n; i++) {
        if (guard && !guard[i]) continue;
        int64_t b = idx[i];
        if (b < 0 || b >= ndst) continue;
        double old = dst[b];
        double nw = vals[i] + old;
        dst[b] = nw;
    }
}

/* Scatter chain (EScatterChain, Lowering.hs:1622-1858): the phases run one
 * after another into the same buffer, each as the (guarded) loop above.
 * Phase k: idx + off[k], vals + off[k], guard + off[k] (or no guard when
 * has_guard[k] == 0), len[k] elements. */
void oracle_scatter_chain_add_f64(const int64_t* idx, const double* vals, const int64_t* guard,
                                  const int64_t* off, const int64_t* len, const int64_t* has_guard,
                                  int64_t nphases, double* dst, int64_t ndst) {
    for (int64_t k = 0; k < nphases; k++)
        oracle_scatter_guarded_add_f64(idx + off[k], vals + off[k], has_guard[k] ? guard + off[k] : NULL,
                                       len[k], dst, ndst);
}

/* The fused weighted histogram (scatter_generate after fusion rule 6,
 * Fusion.hs:2010-2030): route = x % nbins (C truncating %), value = x * x,
 * both recomputed from data[i]; destination fill [nbins] 0; privatized. */
void oracle_weighted_histogram(const int64_t* data, int64_t n, int64_t nbins, int64_t* hist) {
    int nthr = omp_get_max_threads();
    int64_t* grids = (int64_t*)calloc((size_t)nthr * (size_t)nbins, sizeof(int64_t));
    if (grids == NULL) {
        fprintf(stderr, "hydrangea: failed to allocate privatized scatter buffer\n");
        exit(1);
    }
    memset(hist, 0, (size_t)nbins * sizeof(int64_t));
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        int nt = omp_get_num_threads();
        int64_t* priv = grids + (size_t)tid * (size_t)nbins;
#pragma omp for
        for (int64_t i = 0; i < n; i++) {
            int64_t x = data[i];
            int64_t b = x % nbins;
            int64_t v = x * x;
            if (b < 0 || b >= nbins) continue;
            priv[b] += v;
        }
#pragma omp for
        for (int64_t m = 0; m < nbins; m++) {
            int64_t sum = hist[m];
            for (int t = 0; t < nt; t++) sum += grids[(size_t)t * (size_t)nbins + m];
            hist[m] = sum;
        }
    }
    free(grids);
}

/* Same program, serial loop (a different algorithm, used to pin the above). */
void oracle_weighted_histogram_serial(const int64_t* data, int64_t n, int64_t nbins, int64_t* hist) {
    memset(hist, 0, (size_t)nbins * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        int64_t x = data[i];
        int64_t b = x % nbins;
        if (b < 0 || b >= nbins) continue;
        int64_t old = hist[b];
        hist[b] = x * x + old;
    }
}

/* Batched instances: instance k reads data[off[k] .. off[k]+len[k]) into
 * hist[hoff[k] .. hoff[k]+bins[k]). */
void oracle_batched_histograms(const int64_t* data, const int64_t* off, const int64_t* len,
                               const int64_t* bins, const int64_t* hoff, int64_t count, int64_t* hist) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < count; k++) {
        int64_t* h = hist + hoff[k];
        int64_t nb = bins[k];
        memset(h, 0, (size_t)nb * sizeof(int64_t));
        const int64_t* d = data + off[k];
        for (int64_t i = 0; i < len[k]; i++) {
            int64_t x = d[i];
            int64_t b = x % nb;
            if (b < 0 || b >= nb) continue;
            h[b] += x * x;
        }
    }
}

/* ---- reductions ------------------------------------------------------------ */

/* reduce (+.) init (zipwith (*.) xs ys) as emitted with tiling/polyhedral:
 * `#pragma omp parallel for simd simdlen(4) reduction(+:acc)` (README.md:26-32,
 * CodegenC.hs:1306-1413).  Restated deterministically: static schedule over
 * `nthreads` contiguous chunks, 4 SIMD lane partials per chunk (lane = i % 4
 * within the chunk), lanes combined in order, chunks combined in thread order,
 * then added to init.  libgomp combines thread partials in an unspecified
 * order; only the tolerance of the result is part of the contract. */
double oracle_dot(const double* xs, const double* ys, int64_t n, double init, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    double* part = (double*)calloc((size_t)nthreads, sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int t = 0; t < nthreads; t++) {
        int64_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
        double lane[4] = {0.0, 0.0, 0.0, 0.0};
        int64_t i = lo;
        for (; i + 3 < hi; i += 4) {
            for (int l = 0; l < 4; l++) {
                double m = xs[i + l] * ys[i + l];
                lane[l] = lane[l] + m;
            }
        }
        for (int l = 0; i < hi; i++, l++) {
            double m = xs[i] * ys[i];
            lane[l] = lane[l] + m;
        }
        part[t] = ((lane[0] + lane[1]) + lane[2]) + lane[3];
    }
    double acc = init;
    for (int t = 0; t < nthreads; t++) acc = acc + part[t];
    free(part);
    return acc;
}

/* Serial form of the same reduction (Lowering.hs:3517-3535: acc = acc + elem). */
double oracle_dot_serial(const double* xs, const double* ys, int64_t n, double init) {
    double acc = init;
    for (int64_t i = 0; i < n; i++) {
        double m = xs[i] * ys[i];
        acc = acc + m;
    }
    return acc;
}

/* Nearly exact dot: each product split exactly with fma (x*y = p + e), the
 * sum accumulated in double-double (Dekker/Knuth two-sum).  A check of
 * oracle_dot, not a restatement of the reference. */
double oracle_dot_dd(const double* xs, const double* ys, int64_t n, double init) {
    double hi = init, lo = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double p = xs[i] * ys[i];
        double e = fma(xs[i], ys[i], -p);
        double s = hi + p;
        double bp = s - hi;
        double err = (hi - (s - bp)) + (p - bp);
        hi = s;
        lo += err + e;
    }
    return hi + lo;
}

/* ---- map chain ------------------------------------------------------------ */

/* The fused zipwith of programs/map_chain.hyd: one three-address statement
 * per op (CodegenC.hs:2293-2400), abs_f -> fabs, sqrt -> sqrt, expf -> exp
 * (genUnOp, CodegenC.hs:2512-2525), `if a >. b` as an SIf (:978-1001). */
static inline double map_chain_op(double x, double y) {
    double t1 = 2.0 * x;
    double t2 = t1 + y;
    double t3 = fabs(t2);
    double t4 = sqrt(t3);
    double t5 = t4 * x;
    double t6 = t5 - y;
    double t7;
    if (t6 > 0.0) {
        t7 = t6;
    } else {
        t7 = 0.0;
    }
    double t8 = t7 * 1.0e-3;
    return exp(t8);
}

void oracle_map_chain(const double* xs, const double* ys, int64_t n, double* out) {
#pragma omp parallel for simd schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = map_chain_op(xs[i], ys[i]);
}

/* libm exp over an array (the exp the map chain calls). */
void oracle_exp(const double* x, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = exp(x[i]);
}

void oracle_sqrt(const double* x, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = sqrt(x[i]);
}

/* ---- stencils ------------------------------------------------------------- */

/* Bodies in source order (one IEEE op per statement).  body 0: heat step
 * `a 0 0 +. alpha *. (a(-1)0 +. a 1 0 +. a 0(-1) +. a 0 1 -. 4.0 *. a 0 0)`;
 * 1: Laplacian of bench/stencil/stencil_2d.hyd; 2: the Jacobi average of
 * bench/stencil/jacobi_2d.hyd; 3: the 5-point average of examples/heat_2d.hyd. */
static inline double stencil_body(int body, double alpha, double n, double s, double w, double e,
                                  double c) {
    double t1 = n + s;
    double t2 = t1 + w;
    double t3 = t2 + e;
    switch (body) {
        case 0: {
            double t4 = 4.0 * c;
            double t5 = t3 - t4;
            double t6 = alpha * t5;
            return c + t6;
        }
        case 1: {
            double t4 = c * -4.0;
            return t3 + t4;
        }
        case 2: return t3 / 4.0;
        default: {
            double t4 = t3 + c;
            return t4 / 5.0;
        }
    }
}

static inline int64_t clampi(int64_t v, int64_t lo, int64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* stencil clamp over [h, w] (Lowering.hs:1932-1996, clamp index folding
 * :2186-2213): every region of the split-boundary emission evaluates the
 * body on clamped neighbour indices, so one clamped loop nest equals them. */
void oracle_stencil2d_clamp(const double* in, int64_t h, int64_t w, int body, double alpha,
                            double* out) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t i = 0; i < h; i++) {
        for (int64_t j = 0; j < w; j++) {
            double c = in[i * w + j];
            double nn = in[clampi(i - 1, 0, h - 1) * w + j];
            double ss = in[clampi(i + 1, 0, h - 1) * w + j];
            double ww = in[i * w + clampi(j - 1, 0, w - 1)];
            double ee = in[i * w + clampi(j + 1, 0, w - 1)];
            out[i * w + j] = stencil_body(body, alpha, nn, ss, ww, ee, c);
        }
    }
}

/* iterate n init step (Lowering.hs:840-857) with the ping-pong of
 * CFGOpt.hs:824-900: cur = copy(init); loop { next = step(cur); rotate }. */
void oracle_iterate_stencil2d(const double* init, int64_t h, int64_t w, int body, double alpha,
                              int64_t iters, double* out) {
    size_t bytes = (size_t)h * (size_t)w * sizeof(double);
    double* cur = (double*)malloc(bytes ? bytes : 1);
    double* nxt = (double*)malloc(bytes ? bytes : 1);
    memcpy(cur, init, bytes);
    for (int64_t t = 0; t < iters; t++) {
        oracle_stencil2d_clamp(cur, h, w, body, alpha, nxt);
        double* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    memcpy(out, cur, bytes);
    free(cur);
    free(nxt);
}

/* 1-D clamp stencil with a general offset list and integer sum body, for
 * the reference's 1-D known answers (StencilSpec.hs:23-70). */
void oracle_stencil1d_clamp_sum(const int64_t* in, int64_t n, const int64_t* offs, int noffs,
                                int64_t* out) {
    for (int64_t i = 0; i < n; i++) {
        int64_t acc = 0;
        for (int k = 0; k < noffs; k++) acc = acc + in[clampi(i + offs[k], 0, n - 1)];
        out[i] = acc;
    }
}

/* ---- scans, segmented reduce ---------------------------------------------- */

/* scan family (Lowering.hs:3563-3607) with (+): acc = init; for k: ix = k or
 * n-1-k; exclusive writes acc before adding, inclusive after. */
void oracle_scan(const int64_t* in, int64_t n, int64_t init, int inclusive, int reverse, int64_t* out) {
    int64_t acc = init;
    for (int64_t k = 0; k < n; k++) {
        int64_t ix = reverse ? n - 1 - k : k;
        int64_t elem = in[ix];
        if (inclusive) {
            acc = acc + elem;
            out[ix] = acc;
        } else {
            out[ix] = acc;
            acc = acc + elem;
        }
    }
}

/* segmented_reduce (Lowering.hs:874-938): nseg = len(offsets) - 1. */
void oracle_segmented_reduce(const int64_t* offsets, int64_t noffsets, const int64_t* vals,
                             int64_t init, int64_t* out) {
    for (int64_t seg = 0; seg + 1 < noffsets; seg++) {
        int64_t start = offsets[seg], stop = offsets[seg + 1];
        int64_t len = stop - start;
        int64_t acc = init;
        for (int64_t k = 0; k < len; k++) acc = acc + vals[start + k];
        out[seg] = acc;
    }
}

/* ---- sort_indices, gather, COO -> CSR ------------------------------------- */

typedef struct {
    int64_t key;
    int64_t idx;
} entry_t;

/* LSD radix sort of (key, idx) entries (hydrangea_runtime.h:391-426): passes
 * from the largest key (8 bits each, at least one), stable counting sort per
 * pass, result copied back after an odd pass count. */
static void radix_sort_entries(entry_t* a, entry_t* b, int64_t n) {
    int64_t max_key = 0;
    for (int64_t i = 0; i < n; i++)
        if (a[i].key > max_key) max_key = a[i].key;
    int passes = 0;
    int64_t mk = max_key;
    while (mk) {
        passes++;
        mk >>= 8;
    }
    if (passes == 0) passes = 1;
    entry_t *src = a, *dst = b;
    for (int p = 0; p < passes; p++) {
        int shift = p * 8;
        int64_t count[256];
        memset(count, 0, sizeof count);
        for (int64_t i = 0; i < n; i++) count[(src[i].key >> shift) & 0xFF]++;
        int64_t sum = 0;
        for (int j = 0; j < 256; j++) {
            int64_t c = count[j];
            count[j] = sum;
            sum += c;
        }
        for (int64_t i = 0; i < n; i++) {
            int64_t d = (src[i].key >> shift) & 0xFF;
            dst[count[d]++] = src[i];
        }
        entry_t* t = src;
        src = dst;
        dst = t;
    }
    if (passes % 2 == 1) memcpy(a, b, (size_t)n * sizeof(entry_t));
}

/* hyd_sort_indices (hydrangea_runtime.h:428-456). Returns the pass count. */
int oracle_sort_indices(const int64_t* keys, int64_t n, int64_t* perm) {
    entry_t* e = (entry_t*)malloc((size_t)(n ? n : 1) * sizeof(entry_t));
    entry_t* t = (entry_t*)malloc((size_t)(n ? n : 1) * sizeof(entry_t));
    for (int64_t i = 0; i < n; i++) {
        e[i].key = keys[i];
        e[i].idx = i;
    }
    radix_sort_entries(e, t, n);
    for (int64_t i = 0; i < n; i++) perm[i] = e[i].idx;
    int64_t max_key = 0;
    for (int64_t i = 0; i < n; i++)
        if (keys[i] > max_key) max_key = keys[i];
    int passes = 0;
    while (max_key) {
        passes++;
        max_key >>= 8;
    }
    free(e);
    free(t);
    return passes ? passes : 1;
}

/* gather (Lowering.hs:1860-1898): out[i] = src[perm[i]]. */
void oracle_gather(const int64_t* perm, int64_t n, const int64_t* src, int64_t* out) {
    for (int64_t i = 0; i < n; i++) out[i] = src[perm[i]];
}

/* coo_sum_duplicates (Lowering.hs:966-1067), the serial loop with the
 * loop-carried previous row/col/val.  Outputs are `len` long (the shape of
 * the rows input), zero tail.  Returns the merged count. */
int64_t oracle_coo_sum_duplicates(const int64_t* rows, const int64_t* cols, const int64_t* vals,
                                  int64_t nnz, int64_t len, int64_t* out_rows, int64_t* out_cols,
                                  int64_t* out_vals) {
    memset(out_rows, 0, (size_t)len * sizeof(int64_t));
    memset(out_cols, 0, (size_t)len * sizeof(int64_t));
    memset(out_vals, 0, (size_t)len * sizeof(int64_t));
    int64_t out_nnz = 0, prev_row = 0, prev_col = 0, prev_val = 0;
    for (int64_t i = 0; i < nnz; i++) {
        int64_t row = rows[i], col = cols[i], val = vals[i];
        int64_t has_prev = out_nnz > 0;
        if (has_prev && row == prev_row && col == prev_col) {
            int64_t merged = prev_val + val;
            prev_val = merged;
            out_vals[out_nnz - 1] = merged;
        } else {
            out_rows[out_nnz] = row;
            out_cols[out_nnz] = col;
            out_vals[out_nnz] = val;
            prev_row = row;
            prev_col = col;
            prev_val = val;
            out_nnz = out_nnz + 1;
        }
    }
    return out_nnz;
}

/* csr_from_sorted_coo (Lowering.hs:1068-1147). */
void oracle_csr_from_sorted_coo(int64_t nrows, int64_t nnz, const int64_t* rows,
                                const int64_t* cols, const int64_t* vals, int64_t len,
                                int64_t* row_ptr, int64_t* col_idx, int64_t* out_vals) {
    int64_t* counts = (int64_t*)calloc((size_t)(nrows ? nrows : 1), sizeof(int64_t));
    memset(col_idx, 0, (size_t)len * sizeof(int64_t));
    memset(out_vals, 0, (size_t)len * sizeof(int64_t));
    for (int64_t i = 0; i < nnz; i++) {
        int64_t row = rows[i];
        if (row >= 0 && row < nrows) counts[row] = counts[row] + 1;
        col_idx[i] = cols[i];
        out_vals[i] = vals[i];
    }
    int64_t acc = 0;
    for (int64_t i = 0; i < nrows; i++) {
        row_ptr[i] = acc;
        acc = acc + counts[i];
    }
    row_ptr[nrows] = acc;
    free(counts);
}

/* The COO -> CSR program (bench/coo_csr_build/coo_csr_build.hyd:44-62):
 * packed keys, sort_indices, gather, coo_sum_duplicates, CSR.  Returns the
 * merged count. */
int64_t oracle_coo_to_csr(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* rows,
                          const int64_t* cols, const int64_t* vals, int64_t* row_ptr,
                          int64_t* col_idx, int64_t* out_vals) {
    size_t b = (size_t)(nnz ? nnz : 1) * sizeof(int64_t);
    int64_t *keys = malloc(b), *perm = malloc(b), *sr = malloc(b), *sc = malloc(b), *sv = malloc(b);
    int64_t *cr = malloc(b), *cc = malloc(b), *cv = malloc(b);
    for (int64_t i = 0; i < nnz; i++) keys[i] = rows[i] * ncols + cols[i];
    oracle_sort_indices(keys, nnz, perm);
    oracle_gather(perm, nnz, rows, sr);
    oracle_gather(perm, nnz, cols, sc);
    oracle_gather(perm, nnz, vals, sv);
    int64_t m = oracle_coo_sum_duplicates(sr, sc, sv, nnz, nnz, cr, cc, cv);
    oracle_csr_from_sorted_coo(nrows, m, cr, cc, cv, nnz, row_ptr, col_idx, out_vals);
    free(keys);
    free(perm);
    free(sr);
    free(sc);
    free(sv);
    free(cr);
    free(cc);
    free(cv);
    return m;
}

/* ---- CSV text (runtime/hydrangea_runtime.h:549-594, runtime/hyd_write_csv.c:5-59) ----
 * The reference reads element i with fscanf "%lld" / "%lf" after (for i > 0)
 * skipping ',', ' ', '\n', '\r' with fgetc + ungetc, and writes rows with
 * fprintf "%lld" / "%.17g", ',' between elements and '\n' after each row (a
 * 2-D array row by row, anything else as one row).  These restatements run the
 * same loops over in-memory FILEs (fmemopen / open_memstream), so the number
 * conversions are the C library's own, as in the reference.  Where the
 * reference prints a message and exits, the read returns the element index
 * and the kind: 1 = "parse error at element i", 2 = "too few values". */
int oracle_csv_read(const char* text, int64_t nbytes, int64_t count, int is_float, void* out,
                    int64_t* err_elem) {
    static char empty[1];
    FILE* f = fmemopen(nbytes ? (void*)text : (void*)empty, (size_t)nbytes, "r");
    int kind = 0;
    *err_elem = -1;
    if (!f) return -1;
    for (int64_t i = 0; i < count; i++) {
        if (i > 0) {
            int c;
            while ((c = fgetc(f)) != EOF && (c == ',' || c == ' ' || c == '\n' || c == '\r'))
                ;
            if (c == EOF) {
                kind = 2;
                *err_elem = i;
                break;
            }
            ungetc(c, f);
        }
        int r = is_float ? fscanf(f, "%lf", (double*)out + i)
                         : fscanf(f, "%lld", (long long*)((int64_t*)out + i));
        if (r != 1) {
            kind = 1;
            *err_elem = i;
            break;
        }
    }
    fclose(f);
    return kind;
}

/* Writes rows x row_len values; returns a malloc'd buffer (oracle_free). */
char* oracle_csv_write(const void* vals, int64_t rows, int64_t row_len, int is_float, int64_t* nbytes) {
    char* buf = NULL;
    size_t len = 0;
    FILE* f = open_memstream(&buf, &len);
    if (!f) return NULL;
    for (int64_t i = 0; i < rows; i++) {
        for (int64_t j = 0; j < row_len; j++) {
            if (j) fputc(',', f);
            if (is_float) fprintf(f, "%.17g", ((const double*)vals)[i * row_len + j]);
            else fprintf(f, "%lld", (long long)((const int64_t*)vals)[i * row_len + j]);
        }
        fputc('\n', f);
    }
    fclose(f);
    *nbytes = (int64_t)len;
    return buf;
}

void oracle_free(void* p) { free(p); }

/* One fscanf "%lld" / "%lf" per string (strings back to back in text, string
 * k = [offs[k], offs[k+1])): r[k] = fscanf's result, consumed[k] = ftell
 * afterwards, value in out[k]. */
void oracle_scan_tokens(const char* text, const int64_t* offs, int64_t n, int is_float, int32_t* r,
                        void* out, int64_t* consumed) {
    static char empty[1];
    for (int64_t k = 0; k < n; k++) {
        const int64_t len = offs[k + 1] - offs[k];
        FILE* f = fmemopen(len ? (void*)(text + offs[k]) : (void*)empty, (size_t)len, "r");
        if (!f) {
            r[k] = -2;
            continue;
        }
        r[k] = is_float ? fscanf(f, "%lf", (double*)out + k) : fscanf(f, "%lld", (long long*)((int64_t*)out + k));
        consumed[k] = (int64_t)ftell(f);
        fclose(f);
    }
}

/* fprintf "%.17g" / "%lld" text of each value: out + k * stride, length lens[k]. */
void oracle_format_values(const void* vals, int64_t n, int is_float, char* out, int64_t stride, int32_t* lens) {
    for (int64_t k = 0; k < n; k++) {
        char* o = out + k * stride;
        lens[k] = is_float ? snprintf(o, (size_t)stride, "%.17g", ((const double*)vals)[k])
                           : snprintf(o, (size_t)stride, "%lld", (long long)((const int64_t*)vals)[k]);
    }
}
