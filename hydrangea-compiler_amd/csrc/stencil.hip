// stencil.hip - 2-D clamp stencils (fp64) and `iterate` over them.
//
// Reference shapes: stencil lowering src/Language/Hydrangea/Lowering.hs:
// 1932-1996 with split-boundary regions :2562-2661 and per-segment clamp
// folding :2186-2213 (emitted sample: bench/stencil/out_interior.c:47-81).
// Every region evaluates the same body on clamped neighbours, so one
// clamped-index kernel reproduces all of them bit for bit.  `iterate` is the
// serial outer loop of Lowering.hs:840-857 with the ping-pong buffers of
// CFGOpt.hs:824-900: the init array is never written and each step reads the
// previous step's output.
//
// CDNA4 mapping: a wave64 owns 64 consecutive columns (512-byte coalesced
// rows); each thread walks `kRows` rows keeping north/centre/south in
// registers, so each input row is read once per column strip; west/east come
// from the neighbouring lanes through cross-lane moves (the two strip edges
// load their outside neighbour).  Outputs are written with one coalesced
// store per row.  All steps of an iterate are captured once into a hipGraph.
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

#include "hyd_internal.hpp"

namespace hyd {
namespace {

// Bodies: each is the source expression of one program, one IEEE op per
// statement in source order (fp contraction is off).
struct HeatBody {  // a 0 0 +. alpha *. (a(-1)0 +. a 1 0 +. a 0(-1) +. a 0 1 -. 4.0 *. a 0 0)
    double alpha;
    __device__ __forceinline__ double operator()(double n, double s, double w, double e,
                                                 double c) const {
        double t1 = n + s;
        double t2 = t1 + w;
        double t3 = t2 + e;
        double t4 = 4.0 * c;
        double t5 = t3 - t4;
        double t6 = alpha * t5;
        return c + t6;
    }
};
struct LaplacianBody {  // a(-1)0 +. a 1 0 +. a 0(-1) +. a 0 1 +. (a 0 0 *. (-4.0))
    double alpha;
    __device__ __forceinline__ double operator()(double n, double s, double w, double e,
                                                 double c) const {
        double t1 = n + s;
        double t2 = t1 + w;
        double t3 = t2 + e;
        double t4 = c * -4.0;
        return t3 + t4;
    }
};
struct Jacobi4Body {  // (a(-1)0 +. a 1 0 +. a 0(-1) +. a 0 1) /. 4.0
    double alpha;
    __device__ __forceinline__ double operator()(double n, double s, double w, double e,
                                                 double c) const {
        (void)c;
        double t1 = n + s;
        double t2 = t1 + w;
        double t3 = t2 + e;
        return t3 / 4.0;
    }
};
struct Avg5Body {  // (a(-1)0 +. a 1 0 +. a 0(-1) +. a 0 1 +. a 0 0) /. 5.0
    double alpha;
    __device__ __forceinline__ double operator()(double n, double s, double w, double e,
                                                 double c) const {
        double t1 = n + s;
        double t2 = t1 + w;
        double t3 = t2 + e;
        double t4 = t3 + c;
        return t4 / 5.0;
    }
};

constexpr int kWaves = 4;   // waves per workgroup, stacked along rows
constexpr int kRows = 16;   // rows walked by each thread

__device__ __forceinline__ double shfl_up1(double v) {
    return __shfl_up(v, 1, 64);
}
__device__ __forceinline__ double shfl_down1(double v) {
    return __shfl_down(v, 1, 64);
}

template <class Body>
__global__ __launch_bounds__(64 * kWaves) void stencil2d_clamp_kernel(const double* __restrict__ in,
                                                                      int64_t h, int64_t w,
                                                                      Body body,
                                                                      double* __restrict__ out) {
    const int lane = threadIdx.x;  // 0..63
    const int64_t j = (int64_t)blockIdx.x * 64 + lane;
    const int64_t r0 = ((int64_t)blockIdx.y * kWaves + threadIdx.y) * kRows;
    if (r0 >= h) return;  // whole wave leaves together (r0 is wave-uniform)
    const bool active = j < w;
    const int64_t jc = active ? j : w - 1;  // inactive lanes mirror the last column
    const int64_t jw = j - 1 < 0 ? 0 : j - 1;
    const int64_t je = j + 1 >= w ? w - 1 : j + 1;
    const int64_t r_end = r0 + kRows < h ? r0 + kRows : h;

    const double* row_up = in + (r0 > 0 ? r0 - 1 : 0) * w;
    double up = row_up[jc];
    double cur = in[r0 * w + jc];
    for (int64_t r = r0; r < r_end; ++r) {
        const int64_t rd = r + 1 < h ? r + 1 : h - 1;
        double down = in[rd * w + jc];
        // neighbours along the row: lanes hold consecutive columns
        double west = shfl_up1(cur);
        double east = shfl_down1(cur);
        if (lane == 0 || j - 1 < 0 || !active) west = in[r * w + (active ? jw : jc)];
        if (lane == 63 || j + 1 >= w || !active) east = in[r * w + (active ? je : jc)];
        double v = body(up, down, west, east, cur);
        if (active) out[r * w + j] = v;
        up = cur;
        cur = down;
    }
}

// Two columns per lane (16-byte loads and stores): a wave covers 128
// columns; lane l holds columns c0 = 128*bx + 2l and c0+1.  West of c0 is
// lane l-1's second column, east of c0+1 is lane l+1's first column (the
// strip edges load theirs); clamping at the grid edge makes a column its own
// neighbour.  Requires an even width and 16-byte aligned rows.
constexpr int kRows2 = 32;

template <class Body>
__global__ __launch_bounds__(64 * kWaves) void stencil2d_clamp_kernel2(const double* __restrict__ in,
                                                                       int64_t h, int64_t w,
                                                                       Body body,
                                                                       double* __restrict__ out) {
    const int lane = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * 128 + 2 * lane;
    const int64_t r0 = ((int64_t)blockIdx.y * kWaves + threadIdx.y) * kRows2;
    if (r0 >= h) return;  // wave-uniform
    const bool active = c0 < w;
    const int64_t cc = active ? c0 : w - 2;  // inactive lanes mirror the last pair
    const int64_t r_end = r0 + kRows2 < h ? r0 + kRows2 : h;
    const bool left_edge = c0 == 0;
    const bool right_edge = c0 + 2 >= w;

    v2d up = *reinterpret_cast<const v2d*>(in + (r0 > 0 ? r0 - 1 : 0) * w + cc);
    v2d cur = *reinterpret_cast<const v2d*>(in + r0 * w + cc);
    for (int64_t r = r0; r < r_end; ++r) {
        const int64_t rd = r + 1 < h ? r + 1 : h - 1;
        v2d down = *reinterpret_cast<const v2d*>(in + rd * w + cc);
        double from_left = __shfl_up(cur.y, 1, 64);
        double from_right = __shfl_down(cur.x, 1, 64);
        double west_x, east_y;
        if (left_edge)
            west_x = cur.x;
        else if (lane == 0)
            west_x = in[r * w + cc - 1];
        else
            west_x = from_left;
        if (right_edge)
            east_y = cur.y;
        else if (lane == 63)
            east_y = in[r * w + cc + 2];
        else
            east_y = from_right;
        v2d v;
        v.x = body(up.x, down.x, west_x, cur.y, cur.x);
        v.y = body(up.y, down.y, cur.x, east_y, cur.y);
        if (active) st_stream2(out + r * w + c0, v);
        up = cur;
        cur = down;
    }
}

// Whole-wave lane shifts (DPP wave_shr:1 / wave_shl:1, two 32-bit moves per
// double): lane i receives lane i-1's value (from_lower_lane) or lane i+1's
// (from_upper_lane); the lane without a source receives 0 (bound_ctrl).
__device__ __forceinline__ double dpp_shift(double v, int ctrl_is_shr) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    if (ctrl_is_shr) {
        lo = __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, true);
        hi = __builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, true);
    } else {
        lo = __builtin_amdgcn_mov_dpp(lo, 0x130, 0xf, 0xf, true);
        hi = __builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true);
    }
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_lower_lane(double v) { return dpp_shift(v, 1); }
__device__ __forceinline__ double from_upper_lane(double v) { return dpp_shift(v, 0); }

// Temporal blocking: `steps` (<= kTK) stencil steps in one launch.  A
// workgroup owns a kTR x kTC output tile and loads it with a kTK-cell halo:
// a 64-row x 256-column box, lane l holding columns 4l .. 4l+3 of the box,
// each of the 4 waves 16 rows, all in registers.  Four columns per lane need
// one cross-lane move per cell and step (the two outer columns' neighbours),
// against two with two columns per lane, and the wider box recomputes 1.42x
// instead of 1.52x per output cell: the step is VALU-issue bound (7 fp64
// operations per cell fixed by the body).  Measured on MI355X per 8192^2
// step: 35.8 us, against 38.4 us with two columns per lane; taller boxes
// (5, 6 or 8 waves; 20 or 24 rows per wave) and sweeps alternating direction
// (no register copies, but 256+ VGPRs) ran slower, for lack of occupancy.
// Each step recomputes the whole box: north/south come from
// the wave's own registers (the rows just outside its strip are exchanged
// through LDS), west/east from the lane's own columns or the neighbouring
// lane (DPP); clamping applies at the grid edges (EDGE workgroups only -
// interior boxes touch no grid edge).  The box edges go stale by one cell per
// step, but the region still exact after step s is the box shrunk by s cells
// (not shrunk at grid edges), so after kTK steps the output tile holds
// exactly the values of kTK single-step launches: the same operations on the
// same neighbour values, bit for bit.  HBM sees one read of the haloed box
// and one write of the tile per `steps` steps.
constexpr int kTK = 8;                  // fused steps
constexpr int kTRows = 16;              // rows per wave
constexpr int kTWaves = 4;
constexpr int kTCols = 4;               // columns per lane
constexpr int kTBoxR = kTRows * kTWaves;  // 64
constexpr int kTBoxC = 64 * kTCols;       // 256
constexpr int kTR = kTBoxR - 2 * kTK;     // 48 output rows
constexpr int kTC = kTBoxC - 2 * kTK;     // 240 output columns
constexpr int kTBlock = 64 * kTWaves;
// a lane's columns then lie all inside the output tile or all outside it
static_assert(kTK % kTCols == 0 && kTC % kTCols == 0, "halo and tile widths must be whole lanes");

struct LaneRow {  // one lane's columns of one row
    double c[kTCols];
};

template <class Body, bool EDGE>
__device__ __forceinline__ void temporal_box(const double* __restrict__ in, int64_t h, int64_t w,
                                             int steps, Body body, double* __restrict__ out,
                                             int64_t r0, int64_t c0, LaneRow (*halo)[kTWaves][2][64]) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform
    const int64_t c = c0 - kTK + kTCols * lane;               // this lane's first column
    const int64_t rbase = r0 - kTK + (int64_t)wv * kTRows;  // this wave's first row
    LaneRow v[kTRows];
#pragma unroll
    for (int k = 0; k < kTRows; ++k) {
        const int64_t r = rbase + k;
#pragma unroll
        for (int j = 0; j < kTCols; ++j) {
            if (EDGE)
                v[k].c[j] = (r >= 0 && r < h && c + j >= 0 && c + j < w) ? in[r * w + c + j] : 0.0;
            else
                v[k].c[j] = in[r * w + c + j];
        }
    }
    for (int s = 0; s < steps; ++s) {
        const int p = s & 1;
        halo[p][wv][0][lane] = v[0];
        halo[p][wv][1][lane] = v[kTRows - 1];
        __syncthreads();  // parity-double-buffered: one barrier per step
        LaneRow prev = wv > 0 ? halo[p][wv - 1][1][lane] : v[0];  // row above the strip
        const LaneRow dn = wv < kTWaves - 1 ? halo[p][wv + 1][0][lane] : v[kTRows - 1];
#pragma unroll
        for (int k = 0; k < kTRows; ++k) {
            const LaneRow cur = v[k];
            const LaneRow south = k + 1 < kTRows ? v[k + 1] : dn;
            const double from_west = from_lower_lane(cur.c[kTCols - 1]);
            const double from_east = from_upper_lane(cur.c[0]);
            bool top = false, bottom = false;
            if (EDGE) {
                const int64_t r = rbase + k;
                top = r == 0;
                bottom = r == h - 1;
            }
#pragma unroll
            for (int j = 0; j < kTCols; ++j) {
                double west = j == 0 ? from_west : cur.c[j - 1];
                double east = j == kTCols - 1 ? from_east : cur.c[j + 1];
                double nn = prev.c[j], ss = south.c[j];
                if (EDGE) {
                    if (top) nn = cur.c[j];
                    if (bottom) ss = cur.c[j];
                    if (c + j == 0) west = cur.c[j];
                    if (c + j == w - 1) east = cur.c[j];
                }
                v[k].c[j] = body(nn, ss, west, east, cur.c[j]);
            }
            prev = cur;
        }
    }
    // the lane's columns are all inside the output tile or all outside it
    // (c0 - c and kTC are multiples of kTCols)
    const bool col_out = c >= c0 && c < c0 + kTC;
#pragma unroll
    for (int k = 0; k < kTRows; ++k) {
        const int64_t r = rbase + k;
        if (col_out && r >= r0 && r < r0 + kTR && r < h) {
#pragma unroll
            for (int j = 0; j < kTCols; ++j)
                if (!EDGE || c + j < w) out[r * w + c + j] = v[k].c[j];
        }
    }
}

template <class Body>
__global__ __launch_bounds__(kTBlock) void stencil2d_temporal_kernel(const double* __restrict__ in,
                                                                     int64_t h, int64_t w, int steps,
                                                                     Body body,
                                                                     double* __restrict__ out) {
    // [step parity][wave][first/last row of the strip][lane]
    __shared__ LaneRow halo[2][kTWaves][2][64];
    const int64_t r0 = (int64_t)blockIdx.y * kTR, c0 = (int64_t)blockIdx.x * kTC;
    // interior: every box cell is inside the grid and off its edge rows/columns
    const bool interior = r0 - kTK >= 1 && r0 - kTK + kTBoxR <= h - 1 && c0 - kTK >= 1 &&
                          c0 - kTK + kTBoxC <= w - 1;
    if (interior)
        temporal_box<Body, false>(in, h, w, steps, body, out, r0, c0, halo);
    else
        temporal_box<Body, true>(in, h, w, steps, body, out, r0, c0, halo);
}

template <class Body>
int launch_temporal(const double* in, int64_t h, int64_t w, int steps, Body body, double* out,
                    hipStream_t s) {
    int64_t gx = ceil_div(w, kTC), gy = ceil_div(h, kTR);
    if (gx > 0x7fffffff || gy > 65535)
        return set_error(HYD_E_INVALID, "stencil: grid %lld x %lld too large", (long long)h, (long long)w);
    stencil2d_temporal_kernel<Body><<<dim3((unsigned)gx, (unsigned)gy), kTBlock, 0, s>>>(in, h, w, steps,
                                                                                       body, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

template <class Body>
int launch_stencil(const double* in, int64_t h, int64_t w, Body body, double* out, hipStream_t s) {
    const bool pairs = (w % 2 == 0) && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    if (pairs) {
        dim3 block(64, kWaves);
        int64_t gx = ceil_div(w, 128);
        int64_t gy = ceil_div(h, (int64_t)kWaves * kRows2);
        if (gx > 0x7fffffff || gy > 65535)
            return set_error(HYD_E_INVALID, "stencil: grid %lld x %lld too large", (long long)h,
                             (long long)w);
        stencil2d_clamp_kernel2<Body><<<dim3((unsigned)gx, (unsigned)gy), block, 0, s>>>(in, h, w, body, out);
        HYD_CHECK_LAUNCH();
        return HYD_OK;
    }
    dim3 block(64, kWaves);
    int64_t gx = ceil_div(w, 64);
    int64_t gy = ceil_div(h, (int64_t)kWaves * kRows);
    if (gx > 0x7fffffff || gy > 65535)
        return set_error(HYD_E_INVALID, "stencil: grid %lld x %lld too large", (long long)h,
                         (long long)w);
    stencil2d_clamp_kernel<Body><<<dim3((unsigned)gx, (unsigned)gy), block, 0, s>>>(in, h, w, body, out);
    HYD_CHECK_LAUNCH();
    return HYD_OK;
}

int launch_body(const double* in, int64_t h, int64_t w, int body, double alpha, double* out,
                hipStream_t s) {
    switch (body) {
        case HYD_STENCIL_HEAT: return launch_stencil(in, h, w, HeatBody{alpha}, out, s);
        case HYD_STENCIL_LAPLACIAN: return launch_stencil(in, h, w, LaplacianBody{alpha}, out, s);
        case HYD_STENCIL_JACOBI4: return launch_stencil(in, h, w, Jacobi4Body{alpha}, out, s);
        case HYD_STENCIL_AVG5: return launch_stencil(in, h, w, Avg5Body{alpha}, out, s);
        default: return set_error(HYD_E_INVALID, "stencil: unknown body %d", body);
    }
}

// hipGraph cache for iterate: one executable graph per (pointers, shape,
// body, iters, stream).
using GraphKey = std::tuple<const double*, double*, double*, int64_t, int64_t, int, uint64_t,
                            int64_t, hipStream_t>;
std::mutex g_graph_mu;
std::map<GraphKey, hipGraphExec_t> g_graphs;

// `steps` consecutive steps from src into dst (one launch).
int launch_steps(const double* src, int64_t h, int64_t w, int body, double alpha, int steps, double* dst,
                 hipStream_t s) {
    if (steps == 1) return launch_body(src, h, w, body, alpha, dst, s);
    switch (body) {
        case HYD_STENCIL_HEAT: return launch_temporal(src, h, w, steps, HeatBody{alpha}, dst, s);
        case HYD_STENCIL_LAPLACIAN: return launch_temporal(src, h, w, steps, LaplacianBody{alpha}, dst, s);
        case HYD_STENCIL_JACOBI4: return launch_temporal(src, h, w, steps, Jacobi4Body{alpha}, dst, s);
        case HYD_STENCIL_AVG5: return launch_temporal(src, h, w, steps, Avg5Body{alpha}, dst, s);
        default: return set_error(HYD_E_INVALID, "stencil: unknown body %d", body);
    }
}

// Passes of up to kTK fused steps, ping-ponging between buf0 and buf1; the
// result is in the buffer of the last pass.
int64_t num_passes(int64_t iters) { return (iters + kTK - 1) / kTK; }

int record_steps(const double* init, int64_t h, int64_t w, int body, double alpha, int64_t iters,
                 double* buf0, double* buf1, hipStream_t s) {
    const double* src = init;
    int64_t p = 0;
    for (int64_t t = 0; t < iters; t += kTK, ++p) {
        const int steps = (int)std::min<int64_t>(kTK, iters - t);
        double* dst = (p & 1) ? buf1 : buf0;
        HYD_TRY(launch_steps(src, h, w, body, alpha, steps, dst, s));
        src = dst;
    }
    return HYD_OK;
}

}  // namespace
}  // namespace hyd

using namespace hyd;

extern "C" int hydhip_stencil2d_clamp_f64(const double* in, int64_t h, int64_t w, int body,
                                          double alpha, double* out, void* stream) {
    HYD_TRY(ensure_init());
    if (h < 0 || w < 0 || ((h > 0 && w > 0) && (!in || !out)))
        return set_error(HYD_E_INVALID, "hydhip_stencil2d_clamp_f64: bad arguments");
    if (h == 0 || w == 0) return HYD_OK;
    if (in == out) return set_error(HYD_E_INVALID, "hydhip_stencil2d_clamp_f64: in == out");
    return launch_body(in, h, w, body, alpha, out, resolve_stream(stream));
}

extern "C" int hydhip_stencil2d_iterate_f64(const double* init, int64_t h, int64_t w, int body,
                                            double alpha, int64_t iters, double* buf0,
                                            double* buf1, int use_graph, double** result,
                                            void* stream) {
    HYD_TRY(ensure_init());
    if (h < 0 || w < 0 || iters < 0 || !result ||
        ((h > 0 && w > 0) && (!init || !buf0 || (iters > 1 && !buf1))))
        return set_error(HYD_E_INVALID, "hydhip_stencil2d_iterate_f64: bad arguments");
    if (buf0 == init || buf1 == init || (iters > 1 && buf0 == buf1))
        return set_error(HYD_E_INVALID, "hydhip_stencil2d_iterate_f64: buffers must be distinct");
    if (body < HYD_STENCIL_HEAT || body > HYD_STENCIL_AVG5)
        return set_error(HYD_E_INVALID, "hydhip_stencil2d_iterate_f64: unknown body %d", body);
    hipStream_t s = resolve_stream(stream);
    *result = buf0;
    if (h == 0 || w == 0) return HYD_OK;
    const size_t bytes = (size_t)h * (size_t)w * sizeof(double);
    if (iters == 0) {
        HYD_TRY_HIP(hipMemcpyAsync(buf0, init, bytes, hipMemcpyDeviceToDevice, s));
        return HYD_OK;
    }
    *result = ((num_passes(iters) - 1) & 1) ? buf1 : buf0;
    if (!use_graph || iters == 1) return record_steps(init, h, w, body, alpha, iters, buf0, buf1, s);

    uint64_t abits;
    memcpy(&abits, &alpha, sizeof abits);
    GraphKey key{init, buf0, buf1, h, w, body, abits, iters, s};
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        auto it = g_graphs.find(key);
        if (it != g_graphs.end()) exec = it->second;
    }
    if (!exec) {
        hipGraph_t graph = nullptr;
        HYD_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = record_steps(init, h, w, body, alpha, iters, buf0, buf1, s);
        hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc != HYD_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (ec != hipSuccess)
            return set_error(HYD_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess)
            return set_error(HYD_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        std::lock_guard<std::mutex> lk(g_graph_mu);
        g_graphs[key] = exec;
    }
    HYD_TRY_HIP(hipGraphLaunch(exec, s));
    return HYD_OK;
}

// Drops cached graphs that reference `ptr` (called when a buffer is freed).
extern "C" void hydhip_forget_graphs_for(const void* ptr) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (auto it = g_graphs.begin(); it != g_graphs.end();) {
        const GraphKey& k = it->first;
        if ((const void*)std::get<0>(k) == ptr || (const void*)std::get<1>(k) == ptr ||
            (const void*)std::get<2>(k) == ptr) {
            // A launch of this graph may still be queued on its stream; the
            // executable owns the kernel arguments those dispatches read, so
            // it is destroyed only after the stream has drained.
            (void)hipStreamSynchronize(std::get<8>(k));
            (void)hipGraphExecDestroy(it->second);
            it = g_graphs.erase(it);
        } else {
            ++it;
        }
    }
}
