/* The C the HIP backend prints for a program that fills two vectors on the
 * host, runs the map chain and a dot of its result on the device, and prints
 * a checksum (written by hand in the shapes CodegenC prints: no GHC here).
 * It is built with hydhip-cc from this file and map_chain.sites + dot.sites,
 * exactly as HIPBackend.compileAndRunHIP builds a program, and run on the GPU
 * by tests/test_gpu_hydhip_cc.py.
 *
 *   dotmap_program N  ->  prints "dot=<%.17g> bits=<xor of the map chain's bits, hex> threads=<omp threads>"
 *
 * The fill loop stays an OpenMP loop on the host (as CodegenC prints loops
 * the offload pass leaves), so the build must link OpenMP. */
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hyd_hip_runtime.h"

static uint64_t splitmix_at(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int main(int argc, char** argv) {
    int64_t _discarded = 0;
    int64_t n = argc > 1 ? atoll(argv[1]) : 1000;
    hyd_tuple_t shp = (hyd_tuple_t){.ndims = 1, .elems = {(int64_t)n}};
    hyd_array_t* xs = hyd_array_alloc_bytes(shp, sizeof(double));
    hyd_array_t* ys = hyd_array_alloc_bytes(shp, sizeof(double));
    /* a host loop (not offloaded): uniform [-1, 1) from SplitMix64 */
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) {
        ((double*)(void*)xs->data)[i] = 2.0 * ((double)(splitmix_at(17, (uint64_t)i) >> 11) * 0x1.0p-53) - 1.0;
        ((double*)(void*)ys->data)[i] = 2.0 * ((double)(splitmix_at(17, (uint64_t)(n + i)) >> 11) * 0x1.0p-53) - 1.0;
    }
    _discarded = hyd_array_mark_host_dirty(xs);
    _discarded = hyd_array_mark_host_dirty(ys);
    hyd_array_t* out = hyd_array_alloc_device8(shp);
    _discarded = hydk_map_chain_0(n, out, xs, ys);
    double acc = 0.0;
    int64_t red_dim = n;
    acc = hydk_dot_0(acc, red_dim, out, out);
    _discarded = hyd_array_sync_host(out);
    uint64_t bits = 0;
    for (int64_t i = 0; i < n; i++) {
        uint64_t b;
        memcpy(&b, &((double*)(void*)out->data)[i], sizeof b);
        bits ^= b * (uint64_t)(2 * i + 1);
    }
    printf("dot=%.17g bits=%016llx threads=%d\n", acc, (unsigned long long)bits, omp_get_max_threads());
    (void)_discarded;
    hyd_array_free(xs);
    hyd_array_free(ys);
    hyd_array_free(out);
    return 0;
}
