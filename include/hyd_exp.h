/* hyd_exp.h - fp64 exp for the fused map kernels, host and device.
 *
 * The emitted C of the reference lowers the language's `expf` to libm `exp`
 * (src/Language/Hydrangea/CodegenC.hs:2512-2525).  To keep the map chain
 * bit-exact on the GPU, this file implements the published table-driven exp
 * of Szabolcs Nagy (ARM optimized-routines, MIT licence; the algorithm glibc
 * >= 2.28 ships for x86-64 and AArch64):
 *
 *   x = k*ln2/128 + r,  |r| <= ln2/256
 *   exp(x) = 2^(k/128) * exp(r) ~= s + s*(tail + r + r^2 (C2 + r C3) + r^4 (C4 + r C5))
 *
 * with the 2^(k/128) table generated by gen_exp_table.py from exact decimal
 * arithmetic.  Where glibc's FMA build contracts a multiply into an add, the
 * same fused multiply-add is written here explicitly (fma), so the result is
 * the same IEEE operation sequence on the CPU and on gfx950.  Equality with
 * the host libm is checked by the test suite over 10^7+ inputs spanning the
 * whole range (normal, subnormal, overflow, special values).
 *
 * Everything else is plain IEEE binary64 arithmetic; this header must be
 * compiled with -ffp-contract=off.
 */
#ifndef HYD_EXP_H
#define HYD_EXP_H

#include <stdint.h>

#if defined(__HIPCC__)
#define HYD_HD __host__ __device__
#define HYD_FMA __builtin_fma
#else
#include <math.h>
#define HYD_HD
#define HYD_FMA fma
#endif

#if defined(__HIPCC__)
__device__ static const uint64_t hyd_exp_tab_dev[256] = {
#include "hyd_exp_table.inc"
};
#endif
static const uint64_t hyd_exp_tab_host[256] = {
#include "hyd_exp_table.inc"
};

HYD_HD static inline double hyd_as_double(uint64_t u) {
    union { uint64_t u; double d; } c;
    c.u = u;
    return c.d;
}

HYD_HD static inline uint64_t hyd_as_u64(double d) {
    union { uint64_t u; double d; } c;
    c.d = d;
    return c.u;
}

/* |x| >= 512 (or the scale would leave the normal range): rescale to keep the
 * intermediate in range, with the rounding care the subnormal results need. */
HYD_HD static inline double hyd_exp_special(double tmp, uint64_t sbits, uint64_t ki) {
    double scale, y;
    if ((ki & 0x80000000u) == 0) {
        sbits -= 1009ull << 52;
        scale = hyd_as_double(sbits);
        y = 0x1p1009 * HYD_FMA(scale, tmp, scale);
        return y;
    }
    sbits += 1022ull << 52;
    scale = hyd_as_double(sbits);
    double st = scale * tmp;
    y = scale + st;
    if (y < 1.0) {
        double hi, lo;
        lo = scale - y + st;
        hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0) y = 0.0;
    }
    return 0x1p-1022 * y;
}

/* The exp with its table passed in (on the GPU a copy in LDS: the lookup then
 * takes the LDS path instead of competing with a kernel's streaming loads). */
HYD_HD static inline double hyd_exp_with(double x, const uint64_t* tab);

HYD_HD static inline double hyd_exp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return hyd_exp_with(x, hyd_exp_tab_dev);
#else
    return hyd_exp_with(x, hyd_exp_tab_host);
#endif
}

HYD_HD static inline double hyd_exp_with(double x, const uint64_t* tab) {
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128.0;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    const double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double shift = 0x1.8p52;
    const double c2 = 0x1.ffffffffffdbdp-2;
    const double c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5;
    const double c5 = 0x1.1111167a4d017p-7;

    /* 0x3c9 = top12(2^-54), 0x408 = top12(512), 0x409 = top12(1024) */
    const uint32_t abstop = (uint32_t)(hyd_as_u64(x) >> 52) & 0x7ffu;
    /* The main path runs for every input (its integer steps stay in range
     * for any x); the cases it does not cover replace its result afterwards:
     * tiny |x| by a select (no branch: on the GPU about half the lanes of the
     * map chain take it), |x| >= 512, inf and nan by a branch that is rare. */
    double kd = HYD_FMA(inv_ln2_n, x, shift);
    uint64_t ki = hyd_as_u64(kd);
    kd -= shift;
    double r = HYD_FMA(kd, neg_ln2_lo_n, HYD_FMA(kd, neg_ln2_hi_n, x));
    uint64_t idx = 2 * (ki % 128);
    uint64_t top = ki << 45;
    double tail = hyd_as_double(tab[idx]);
    uint64_t sbits = tab[idx + 1] + top;
    double r2 = r * r;
    double tmp = HYD_FMA(r2 * r2, HYD_FMA(r, c5, c4), HYD_FMA(r2, HYD_FMA(r, c3, c2), tail + r));
    double scale = hyd_as_double(sbits);
    double y = HYD_FMA(scale, tmp, scale);
    const double tiny = 1.0 + x;
    y = abstop < 0x3c9u ? tiny : y; /* tiny |x| */
    if (abstop >= 0x408u) {
        if (abstop == 0x408u) return hyd_exp_special(tmp, sbits, ki); /* 512 <= |x| < 1024 */
        if (hyd_as_u64(x) == 0xfff0000000000000ull) return 0.0;      /* -inf */
        if (abstop >= 0x7ffu) return 1.0 + x;                       /* +inf, nan */
        if (hyd_as_u64(x) >> 63) return 0.0;                         /* underflow */
        return hyd_as_double(0x7ff0000000000000ull);                 /* overflow */
    }
    return y;
}

#endif /* HYD_EXP_H */
