// tune.hip - autotuner for the batched scatter: per (bin class, size class)
// it times the LDS-private and the global-atomic form of the batched kernel
// on synthetic instances of that class and keeps the faster one.  This
// replaces the reference's fixed choices (privatize when the destination has
// <= 262144 elements, Parallelize.hs:466-503; fixed 32x32 polyhedral tiles,
// Polyhedral.hs:200-204) with measurements on the device at hand.
//
// It also tunes launch shapes: the streaming grid (workgroups per CU) of the
// map chain and the copy per size class, where results do not depend on the
// grid.
//
// Table file (text, one class per line: "bin_log2 n_log2 mode", or for a
// launch shape "-1-kind n_log2 workgroups_per_cu"), default $HYD_TUNE_CACHE or
// $HOME/.cache/hydhip/tune_<arch>.txt.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "hyd_internal.hpp"

namespace hyd {

int batched_histogram(const int64_t* data, const int64_t* off, const int64_t* len, const int64_t* bins,
                      const int64_t* hoff, int64_t count, int64_t* hist, int policy, hipStream_t s,
                      int64_t* out_tiles);

namespace {

std::mutex g_mu;
std::map<std::pair<int, int>, int> g_table;
bool g_loaded = false;

// ceil / floor of log2 (0 for v <= 1)
int ilog2_ceil(int64_t v) { return v <= 1 ? 0 : 64 - __builtin_clzll((unsigned long long)(v - 1)); }
int ilog2_floor(int64_t v) { return v <= 1 ? 0 : 63 - __builtin_clzll((unsigned long long)v); }

std::string default_path() {
    const char* e = getenv("HYD_TUNE_CACHE");
    if (e && *e) return e;
    const char* home = getenv("HOME");
    std::string dir = std::string(home ? home : "/tmp") + "/.cache/hydhip";
    return dir + "/tune_" + device_info().name + ".txt";
}

void mkdirs(const std::string& path) {
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
}

int load_locked(const std::string& path) {
    FILE* f = fopen(path.c_str(), "r");
    g_loaded = true;
    if (!f) return 0;
    int a, b, m, n = 0;
    while (fscanf(f, "%d %d %d", &a, &b, &m) == 3) {
        g_table[{a, b}] = m;
        ++n;
    }
    fclose(f);
    return n;
}

}  // namespace


int tune_class_bins(int64_t bins) { return ilog2_ceil(bins); }
int tune_class_n(int64_t n) { return ilog2_floor(std::max<int64_t>(n, 1)); }

// Launch-shape entries share the table: key (-1 - kind, floor log2 n), value
// the streaming grid's workgroups per CU.  A call of size n takes the entry
// of the largest tuned size class <= n's; without one, `dflt`.
int tuned_per_cu(int kind, int64_t n, int dflt) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_loaded) load_locked(default_path());
    auto it = g_table.upper_bound({-1 - kind, ilog2_floor(std::max<int64_t>(n, 1))});
    if (it == g_table.begin()) return dflt;
    --it;
    if (it->first.first != -1 - kind || it->second < 1) return dflt;
    return it->second;
}

// The table as a dense array (-1: untuned), under one lock.
void tuned_batch_snapshot(TuneSnapshot* out) {
    memset(out->mode, -1, sizeof out->mode);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_loaded) load_locked(default_path());
    for (auto& kv : g_table)
        if (kv.first.first >= 0 && kv.first.first < 64 && kv.first.second >= 0 && kv.first.second < 64)
            out->mode[kv.first.first][kv.first.second] = (int8_t)kv.second;
}

}  // namespace hyd

using namespace hyd;

extern "C" {

// Sets one class by hand (mode 0 = LDS, 1 = global; -1 removes it).
int hydhip_tune_set(int64_t bins, int64_t n, int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_loaded = true;
    auto key = std::make_pair(ilog2_ceil(bins), ilog2_floor(std::max<int64_t>(n, 1)));
    if (mode < 0)
        g_table.erase(key);
    else
        g_table[key] = mode;
    return HYD_OK;
}

// Sets the streaming grid of one launch-shape class by hand (per_cu <= 0
// removes it).
int hydhip_tune_shape_set(int kind, int64_t n, int per_cu) {
    if (kind < 0 || kind >= HYD_SHAPE_KINDS || n < 1)
        return set_error(HYD_E_INVALID, "hydhip_tune_shape_set: bad kind %d or size %lld", kind, (long long)n);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_loaded) load_locked(default_path());
    auto key = std::make_pair(-1 - kind, ilog2_floor(n));
    if (per_cu <= 0)
        g_table.erase(key);
    else
        g_table[key] = per_cu;
    return HYD_OK;
}

int hydhip_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    g_loaded = true;
    return HYD_OK;
}

static int tune_impl(const char* path, void* stream, int* measured);

// Measures every (bins, size) class of the batched workload and stores the
// table (path NULL -> default).  Returns the number of classes measured, or
// minus a status code.
int hydhip_tune_batched(const char* path, void* stream) {
    int measured = 0;
    int rc = tune_impl(path, stream, &measured);
    return rc == HYD_OK ? measured : -rc;
}

constexpr float kGlobalMargin = 0.9f;  // global must be 10 % faster than LDS

static int tune_impl(const char* path, void* stream, int* measured) {
    HYD_TRY(ensure_init());
    hipStream_t s = resolve_stream(stream);
    const int64_t bin_classes[] = {64, 256, 4096, 65536};
    const int64_t hist_words = (int64_t)1 << 26;  // 512 MiB of output per measurement
    const int n_logs[] = {10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20};
    const int64_t total = (int64_t)1 << 23;  // elements per measurement batch
    int64_t* data = nullptr;
    int64_t* hist = nullptr;
    if (hipMalloc(&data, total * 8) != hipSuccess)
        return set_error(HYD_E_OOM, "tune: allocation failed");
    HYD_TRY(hydhip_fill_splitmix_i64(data, total, 0x74756e65ULL, 0, 33, s));
    hipEvent_t e0, e1;
    HYD_TRY_HIP(hipEventCreate(&e0));
    HYD_TRY_HIP(hipEventCreate(&e1));
    std::map<std::pair<int, int>, int> found;
    for (int64_t nb : bin_classes) {
        for (int nl : n_logs) {
            int64_t n = (int64_t)1 << nl;
            int64_t count = std::max<int64_t>(1, std::min<int64_t>(total / n, hist_words / nb));
            std::vector<int64_t> off(count), len(count, n), bins(count, nb), hoff(count);
            for (int64_t k = 0; k < count; ++k) {
                off[k] = k * n;
                hoff[k] = k * nb;
            }
            if (!hist) {
                if (hipMalloc(&hist, (size_t)hist_words * 8) != hipSuccess) {
                    (void)hipFree(data);
                    return set_error(HYD_E_OOM, "tune: allocation failed");
                }
            }
            float best[2] = {1e30f, 1e30f};
            for (int mode = 0; mode < 2; ++mode) {
                for (int rep = 0; rep < 4; ++rep) {
                    HYD_TRY_HIP(hipEventRecord(e0, s));
                    HYD_TRY(batched_histogram(data, off.data(), len.data(), bins.data(), hoff.data(),
                                              count, hist, mode == 0 ? 1 : 2, s, nullptr));
                    HYD_TRY_HIP(hipEventRecord(e1, s));
                    HYD_TRY_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    HYD_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (rep > 0 && ms < best[mode]) best[mode] = ms;
                }
            }
            // the global form only where it wins clearly: in a mixed batch its
            // atomics share L2 and HBM with the streaming LDS tiles, which an
            // isolated measurement does not see
            found[{ilog2_ceil(nb), nl}] = best[1] < kGlobalMargin * best[0] ? 1 : 0;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(data);
    (void)hipFree(hist);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : found) g_table[kv.first] = kv.second;
    g_loaded = true;
    std::string p = path ? std::string(path) : default_path();
    mkdirs(p);
    FILE* f = fopen(p.c_str(), "w");
    if (f) {
        for (auto& kv : g_table) fprintf(f, "%d %d %d\n", kv.first.first, kv.first.second, kv.second);
        fclose(f);
    }
    *measured = (int)found.size();
    return HYD_OK;
}

// Copies the batched-scatter classes into out[3*i..] (bin_log2, n_log2,
// mode); returns entries.
int hydhip_tune_table(int* out, int max_entries) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_loaded) load_locked(default_path());
    int i = 0;
    for (auto& kv : g_table) {
        if (kv.first.first < 0) continue;  // launch shapes
        if (i >= max_entries) break;
        out[3 * i] = kv.first.first;
        out[3 * i + 1] = kv.first.second;
        out[3 * i + 2] = kv.second;
        ++i;
    }
    return i;
}

// Copies the launch-shape classes into out[3*i..] (kind, n_log2, per_cu).
int hydhip_tune_shape_table(int* out, int max_entries) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_loaded) load_locked(default_path());
    int i = 0;
    for (auto& kv : g_table) {
        if (kv.first.first >= 0) continue;
        if (i >= max_entries) break;
        out[3 * i] = -1 - kv.first.first;
        out[3 * i + 1] = kv.first.second;
        out[3 * i + 2] = kv.second;
        ++i;
    }
    return i;
}

// Measures the streaming grid (1, 2, 4 or 8 workgroups per CU) of the map
// chain and the copy at 2^20, 2^23 and 2^26 elements, keeps the fastest of
// each (best of 3 after a warm-up) and stores the table like
// hydhip_tune_batched.  The dot is not tuned: its partial count fixes the
// order of its sum, so its grid stays a fixed function of n and the device.
// Returns the number of classes measured, or minus a status code.
int hydhip_tune_shapes(const char* path, void* stream) {
    int rc = ensure_init();
    if (rc) return -rc;
    hipStream_t s = resolve_stream(stream);
    const int logs[] = {20, 23, 26};
    const int per_cu[] = {HYD_SHAPE_FLAT, 1, 2, 4, 8};
    const int64_t nmax = (int64_t)1 << 26;
    double* buf = nullptr;
    if (hipMalloc(&buf, (size_t)nmax * 3 * 8) != hipSuccess) {
        set_error(HYD_E_OOM, "tune shapes: allocation failed");
        return -HYD_E_OOM;
    }
    double *xs = buf, *ys = buf + nmax, *out = buf + 2 * nmax;
    rc = hydhip_fill_splitmix_f64(xs, 2 * nmax, 0x73686170ULL, 0, 2.0, -1.0, s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = set_error(HYD_E_HIP, "tune shapes: event creation failed");
    int measured = 0;
    for (int kind = 0; kind < HYD_SHAPE_KINDS && !rc; ++kind) {
        for (int nl : logs) {
            const int64_t n = (int64_t)1 << nl;
            float best = 1e30f;
            int pick = 1;
            for (int pc : per_cu) {
                if ((rc = hydhip_tune_shape_set(kind, n, pc))) break;
                float tbest = 1e30f;
                for (int rep = 0; rep < 4 && !rc; ++rep) {
                    if (hipEventRecord(e0, s) != hipSuccess) rc = set_error(HYD_E_HIP, "tune shapes: event");
                    if (!rc) rc = kind == HYD_SHAPE_MAP_CHAIN ? hydhip_map_chain_f64(xs, ys, n, out, s)
                                                             : hydhip_copy_f64(xs, n, out, s);
                    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                        rc = set_error(HYD_E_HIP, "tune shapes: event");
                    float ms = 0;
                    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = set_error(HYD_E_HIP, "tune shapes: event");
                    if (!rc && rep > 0 && ms < tbest) tbest = ms;
                }
                if (rc) break;
                if (tbest < best) {
                    best = tbest;
                    pick = pc;
                }
            }
            if (rc) break;
            rc = hydhip_tune_shape_set(kind, n, pick);
            ++measured;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (rc) return -rc;
    std::lock_guard<std::mutex> lk(g_mu);
    std::string p = path ? std::string(path) : default_path();
    mkdirs(p);
    FILE* f = fopen(p.c_str(), "w");
    if (f) {
        for (auto& kv : g_table) fprintf(f, "%d %d %d\n", kv.first.first, kv.first.second, kv.second);
        fclose(f);
    }
    return measured;
}

}  // extern "C"
