// csv_num.h - number text <-> int64 / binary64 for the CSV kernels (host and
// device code: the host build serves the library's CPU test hooks).
//
// Reference semantics (runtime/hydrangea_runtime.h:549-594,
// runtime/hyd_write_csv.c:5-59): elements are read with fscanf "%lld" /
// "%lf" and written with fprintf "%lld" / "%.17g" of the C library the
// reference links, glibc.  What that means for the text, restated:
//
//  * "%lld": optional sign, then one or more decimal digits, consumed
//    greedily; a value outside int64 saturates to INT64_MIN / INT64_MAX.
//  * "%lf": glibc's scanf collects a buffer with this state machine and then
//    converts the buffer's longest valid prefix with strtod (correctly
//    rounded, ties to even):
//      - optional sign; "nan" (any case: the buffer stops there, a "(...)"
//        is not read); "inf" or "infinity" (any case; "infi".."infinit" not
//        followed by the rest is a failure);
//      - "0x" / "0X" starts a hexadecimal number, whose exponent marker is p;
//      - then, in order: a digit (or a hex digit before the exponent, in hex
//        mode) is taken; a sign directly after the exponent marker is taken;
//        the exponent marker is taken once a mantissa digit has been seen;
//        one '.' is taken before the exponent; anything else ends the buffer;
//      - a buffer of only the sign, or only the sign and "0x", fails, and so
//        does one with no mantissa digit ("." or "-."), except that "0x."
//        converts as 0 (strtod reads its "0");
//      - an exponent marker (and sign) with no digits after it is consumed
//        but ignored: "1e," reads 1 and leaves ",".
//  * "%.17g": 17 significant digits, correctly rounded (ties to even); the
//    %e form when the decimal exponent X is < -4 or >= 17, else the %f form;
//    trailing zeros of the fraction and a bare '.' removed; exponent with a
//    sign and at least two digits; "inf", "nan" with a '-' when the sign bit
//    is set.
//
// Decimal -> binary64 (scan + dec_to_bits): the first 19 significant digits
// w and an exponent q with x ~ w * 10^q; w is normalised to 64 bits and
// multiplied by a 128-bit scaled power of five (rounded up unless exact), so
// the 192-bit product P over-estimates x by less than one unit of w.  Unless
// a rounding midpoint falls in that window the rounding is decided from P;
// otherwise (and when more than 19 digits straddle a midpoint) the exact path
// compares x with the midpoint in big-integer arithmetic.  The method is the
// Eisel-Lemire algorithm (D. Lemire, "Number parsing at a gigabyte per
// second", Software: Practice and Experience 51(8), 2021) with an exact
// big-integer comparison for the undecided cases.
//
// binary64 -> 17 digits (f64_to_dec17): N = round(|v| * 10^(16 - X)) from the
// same table and the same error window; undecided cases (ties included) are
// settled exactly by the big-integer comparison.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HYD_HD __host__ __device__ __forceinline__

namespace hyd {
namespace csv {

constexpr int kQMin = -342;
constexpr int kQMax = 341;
constexpr int kExactMax = 55;  // 5^q < 2^128: the table entry is exact

__device__ static const uint64_t pow5_dev[kQMax - kQMin + 1][3] = {
#include "hyd_pow5_table.inc"
};
static const uint64_t pow5_host[kQMax - kQMin + 1][3] = {
#include "hyd_pow5_table.inc"
};

HYD_HD const uint64_t* pow5_row(int q) {
#if defined(__HIP_DEVICE_COMPILE__)
    return pow5_dev[q - kQMin];
#else
    return pow5_host[q - kQMin];
#endif
}

constexpr uint64_t kInfBits = 0x7FF0000000000000ULL;
constexpr uint64_t kNanBits = 0x7FF8000000000000ULL;
constexpr uint64_t kSignBit = 0x8000000000000000ULL;
constexpr uint64_t kTen17 = 100000000000000000ULL;
constexpr uint64_t kTen16 = 10000000000000000ULL;

// ---- character classes (c == -1 is the end of the text) -------------------
HYD_HD bool is_digit(int c) { return (unsigned)(c - '0') < 10u; }
// isspace() of the C locale: ' ', \t, \n, \v, \f, \r
HYD_HD bool is_ws(int c) { return c == ' ' || (unsigned)(c - '\t') < 5u; }
// the reference's separator loop: ',', ' ', '\n', '\r'
HYD_HD bool is_sep(int c) { return c == ',' || c == ' ' || c == '\n' || c == '\r'; }
// characters that can lie between two numbers
HYD_HD bool is_gap(int c) { return c == ',' || is_ws(c); }
HYD_HD int lower(int c) { return c | 0x20; }  // exact for letters; never maps a non-letter onto one
HYD_HD bool is_hex_alpha(int c) { return (unsigned)(lower(c) - 'a') < 6u; }

// ---- 192-bit helpers ------------------------------------------------------
struct U192 {
    uint64_t w0, w1, w2;  // w0 least significant
};

HYD_HD U192 mul_64x128(uint64_t a, uint64_t hi, uint64_t lo) {
    const unsigned __int128 pl = (unsigned __int128)a * lo;
    const unsigned __int128 ph = (unsigned __int128)a * hi;
    const unsigned __int128 mid = (pl >> 64) + (uint64_t)ph;
    U192 r;
    r.w0 = (uint64_t)pl;
    r.w1 = (uint64_t)mid;
    r.w2 = (uint64_t)(ph >> 64) + (uint64_t)(mid >> 64);
    return r;
}

HYD_HD bool bit_at(const U192& x, int j) {  // 0 <= j
    if (j >= 192) return false;
    const uint64_t w = j < 64 ? x.w0 : (j < 128 ? x.w1 : x.w2);
    return (w >> (j & 63)) & 1;
}

HYD_HD uint64_t word_mask(int s, int e) {  // bits [s, e) of a word, 0 <= s <= e <= 64
    if (s >= e) return 0;
    const uint64_t hi = e == 64 ? ~0ULL : ((1ULL << e) - 1);
    return hi & ~((1ULL << s) - 1);
}

// bits [a, b) of x all zero
HYD_HD bool bits_zero(const U192& x, int a, int b) {
    if (b > 192) b = 192;
    if (a < 0) a = 0;
    if (a >= b) return true;
    auto clampw = [](int v, int lo) { return v < lo ? 0 : (v > lo + 64 ? 64 : v - lo); };
    const uint64_t m0 = word_mask(clampw(a, 0), clampw(b, 0));
    const uint64_t m1 = word_mask(clampw(a, 64), clampw(b, 64));
    const uint64_t m2 = word_mask(clampw(a, 128), clampw(b, 128));
    return ((x.w0 & m0) | (x.w1 & m1) | (x.w2 & m2)) == 0;
}

// x >> k for 64 <= k (result truncated to 64 bits)
HYD_HD uint64_t shr_192(const U192& x, int k) {
    if (k >= 192) return 0;
    if (k >= 128) return x.w2 >> (k - 128);
    if (k == 64) return x.w1;
    return (x.w1 >> (k - 64)) | (x.w2 << (128 - k));
}

HYD_HD int clz64(uint64_t v) { return __builtin_clzll(v); }

// Encodes mant * 2^E (mant < 2^54 after rounding; E >= -1074) as binary64
// magnitude bits.
HYD_HD uint64_t encode(uint64_t mant, int64_t E) {
    if (mant >> 53) {
        mant >>= 1;
        E += 1;
    }
    if (mant >> 52) {
        const int64_t biased = E + 1075;
        if (biased >= 2047) return kInfBits;
        return ((uint64_t)biased << 52) | (mant & ((1ULL << 52) - 1));
    }
    return mant;  // subnormal (E == -1074) or zero
}

// w * 10^q (w >= 1) rounded to nearest-even, as magnitude bits.  Returns
// false when a midpoint may lie in the product's error window: the exact path
// decides those.
HYD_HD bool dec_to_bits(uint64_t w, int64_t q, uint64_t* out) {
    if (q > 308) {
        *out = kInfBits;
        return true;
    }
    if (q < kQMin) {
        *out = 0;
        return true;
    }
    const uint64_t* row = pow5_row((int)q);
    const int lz = clz64(w);
    const uint64_t wn = w << lz;
    const U192 P = mul_64x128(wn, row[0], row[1]);
    const int64_t f = (int64_t)row[2];
    const int msb = (P.w2 >> 63) ? 191 : 190;
    const int64_t s0 = f + q - lz;  // x ~ P * 2^s0
    int64_t k = msb - 52;           // position of the mantissa's last bit in P
    if (k + s0 < -1074) k = -1074 - s0;
    if (k > 192) {
        *out = 0;
        return true;
    }
    const int kk = (int)k;
    uint64_t mant = shr_192(P, kk);
    const bool rb = bit_at(P, kk - 1);
    bool up, decided = true;
    if (q >= 0 && q <= kExactMax) {
        up = rb && (!bits_zero(P, 0, kk - 1) || (mant & 1));
    } else {
        // P - x lies in (0, wn): a midpoint at P - R (R = bits below the round
        // bit) is inside the window when R < wn
        decided = !(rb && bits_zero(P, 64, kk - 1) && P.w0 < wn);
        up = rb;
    }
    mant += up ? 1 : 0;
    *out = encode(mant, k + s0);  // within one unit of the answer when undecided
    return decided;
}

// hexadecimal w * 2^q (+ nonzero digits below w when sticky), rounded to
// nearest-even, as magnitude bits (exact: no table).
HYD_HD uint64_t hex_to_bits(uint64_t w, int64_t q, bool sticky) {
    if (w == 0) return 0;
    const int lz = clz64(w);
    const int64_t e2 = q + 63 - lz;  // exponent of the leading bit
    if (e2 > 1023) return kInfBits;
    if (e2 < -1076) return 0;  // below 2^-1075, half the smallest subnormal
    int64_t lsb = e2 - 52;
    if (lsb < -1074) lsb = -1074;
    const int64_t sh = lsb - q;  // bits of w below the mantissa's last bit
    uint64_t mant;
    if (sh <= 0) {
        mant = w << (-sh);
    } else {
        mant = sh >= 64 ? 0 : (w >> sh);
        const bool rb = sh <= 64 ? ((w >> (sh - 1)) & 1) : false;
        const bool rest = (sh >= 2 ? (w & (sh - 1 >= 64 ? ~0ULL : ((1ULL << (sh - 1)) - 1))) != 0 : false) ||
                          sticky;
        if (rb && (rest || (mant & 1))) mant += 1;
    }
    return encode(mant, lsb);
}

// ---- token scanners --------------------------------------------------------
// At: callable int(int64_t position) returning the byte there or -1 past the end.

// "%lld" at p: consumed length (> 0), or -1 when no number can be read.
template <class At>
HYD_HD int64_t scan_i64(const At& at, int64_t p, int64_t* value) {
    int64_t i = p;
    int c = at(i);
    bool neg = false;
    if (c == '+' || c == '-') {
        neg = c == '-';
        c = at(++i);
    }
    if (!is_digit(c)) return -1;
    uint64_t acc = 0;
    bool ovf = false;
    do {
        const uint64_t d = (uint64_t)(c - '0');
        if (!ovf) {
            if (acc > (0xFFFFFFFFFFFFFFFFULL - d) / 10) ovf = true;
            else acc = acc * 10 + d;
        }
        c = at(++i);
    } while (is_digit(c));
    if (value) {
        if (neg) *value = (ovf || acc > 0x8000000000000000ULL) ? (int64_t)0x8000000000000000ULL
                                                               : (int64_t)(0 - acc);
        else *value = (ovf || acc > 0x7FFFFFFFFFFFFFFFULL) ? 0x7FFFFFFFFFFFFFFFLL : (int64_t)acc;
    }
    return i - p;
}

enum FloatKind : int { kDec = 0, kHex = 1, kInf = 2, kNan = 3, kHexZero = 4 };

struct FloatTok {
    int64_t len;     // consumed characters, -1: conversion failure
    int kind;
    bool neg;
    uint64_t w;      // leading significant digits (19 decimal / 16 hex)
    int64_t q;       // x = w * 10^q (decimal) or w * 2^q (hex), before dropped digits
    bool truncated;  // nonzero digits dropped after those in w
};

// Digit sink keeping the first kMax significant digits of the mantissa.
struct FastSink {
    static constexpr int kMaxDec = 19, kMaxHex = 16;
    uint64_t w = 0;
    int nsig = 0;
    int64_t qadj = 0;
    bool truncated = false;
    HYD_HD void digit(int d, bool after_point, bool hex) {
        // leading zeros only move the point; the first kMax significant
        // digits build w; later ones only mark truncation (and, before the
        // point, scale w)
        const int unit = hex ? 4 : 1;
        const bool lead = nsig == 0 && d == 0;
        const bool keep = !lead && nsig < (hex ? kMaxHex : kMaxDec);
        const uint64_t wk = hex ? ((w << 4) | (uint64_t)d) : (w * 10 + (uint64_t)d);
        w = keep ? wk : w;
        nsig += keep ? 1 : 0;
        truncated = truncated || (!lead && !keep && d != 0);
        qadj += after_point ? (lead || keep ? -unit : 0) : (!lead && !keep ? unit : 0);
    }
};

constexpr int64_t kExpCap = (int64_t)1 << 50;

// "%lf" at p (no leading whitespace skip: the caller has done it).
template <class At, class Sink>
HYD_HD FloatTok scan_f64(const At& at, int64_t p, Sink& sink) {
    FloatTok r;
    r.len = -1;
    r.kind = kDec;
    r.neg = false;
    r.w = 0;
    r.q = 0;
    r.truncated = false;
    int64_t i = p;
    int c = at(i);
    int sign_len = 0;
    if (c == '+' || c == '-') {
        r.neg = c == '-';
        sign_len = 1;
        c = at(++i);
        if (c < 0) return r;
    }
    if (lower(c) == 'n') {
        if (lower(at(i + 1)) != 'a' || lower(at(i + 2)) != 'n') return r;
        r.kind = kNan;
        r.len = i + 3 - p;
        return r;
    }
    if (lower(c) == 'i') {
        if (lower(at(i + 1)) != 'n' || lower(at(i + 2)) != 'f') return r;
        i += 3;
        if (lower(at(i)) == 'i') {
            if (lower(at(i + 1)) != 'n' || lower(at(i + 2)) != 'i' || lower(at(i + 3)) != 't' ||
                lower(at(i + 4)) != 'y')
                return r;
            i += 5;
        }
        r.kind = kInf;
        r.len = i - p;
        return r;
    }
    bool hex = false, mant_digit = false;
    if (c == '0') {
        c = at(++i);
        if (lower(c) == 'x') {
            hex = true;
            c = at(++i);
        } else {
            mant_digit = true;
            sink.digit(0, false, false);
        }
    }
    // the buffer glibc's loop collects, in its three possible parts:
    // mantissa digits, '.' and digits, then (after a mantissa digit) the
    // exponent marker, one sign and decimal digits
    int64_t expv = 0;
    bool eneg = false, exp_digit = false;
    for (bool after_point = false;; after_point = true) {
        for (;;) {
            int d;
            if (is_digit(c)) d = c - '0';
            else if (hex && is_hex_alpha(c)) d = lower(c) - 'a' + 10;
            else break;
            sink.digit(d, after_point, hex);
            mant_digit = true;
            c = at(++i);
        }
        if (after_point || c != '.') break;
        c = at(++i);
    }
    if (mant_digit && lower(c) == (hex ? 'p' : 'e')) {
        c = at(++i);
        if (c == '+' || c == '-') {
            eneg = c == '-';
            c = at(++i);
        }
        while (is_digit(c)) {
            expv = expv > kExpCap / 10 ? kExpCap : expv * 10 + (c - '0');
            if (expv > kExpCap) expv = kExpCap;
            exp_digit = true;
            c = at(++i);
        }
    }
    const int64_t size = i - p;
    if (size == sign_len) return r;
    if (hex && size == 2 + sign_len) return r;
    if (!mant_digit) {
        if (!hex) return r;  // "." or "-.": strtod reads nothing
        r.kind = kHexZero;   // "0x.": strtod reads the "0"
        r.len = size;
        return r;
    }
    r.kind = hex ? kHex : kDec;
    r.len = size;
    const int64_t e = exp_digit ? (eneg ? -expv : expv) : 0;
    r.q = e + sink.qadj;
    return r;
}

template <class At>
HYD_HD FloatTok scan_f64_fast(const At& at, int64_t p) {
    FastSink sink;
    FloatTok r = scan_f64(at, p, sink);
    r.w = sink.w;
    r.truncated = sink.truncated;
    return r;
}

// ---- the common decimal shape, 8 characters at a time -------------------------
// [sign] digits [. digits] [e|E [sign] digits], found with byte masks on
// 64-bit words (byte 0 = the first character) instead of one character per
// step.  Same result as scan_f64 + FastSink; anything else (hex, inf, nan, a
// lone '.', runs reaching past kSwarSpan bytes) is left to scan_f64.

constexpr int kSwarSpan = 48;  // bytes from the number's start the reader may touch

// bit 7 of each byte -> bit i for byte i (distinct shifted copies: no carries)
HYD_HD uint32_t byte_mask(uint64_t f80) {
    return (uint32_t)(((f80 & 0x8080808080808080ULL) * 0x0002040810204081ULL) >> 56);
}
// bit 7 set in exactly the zero bytes of x
HYD_HD uint64_t zero_bytes(uint64_t x) {
    return ~(((x & 0x7F7F7F7F7F7F7F7FULL) + 0x7F7F7F7F7F7F7F7FULL) | x) & 0x8080808080808080ULL;
}
HYD_HD uint32_t digit_mask(uint64_t x) {
    const uint64_t hi = (x & 0xF0F0F0F0F0F0F0F0ULL) ^ 0x3030303030303030ULL;  // 0 iff high nibble 3
    const uint64_t lo = ((x & 0x0F0F0F0F0F0F0F0FULL) + 0x0606060606060606ULL) & 0xF0F0F0F0F0F0F0F0ULL;  // 0 iff low <= 9
    return byte_mask(zero_bytes(hi | lo));
}
HYD_HD uint32_t eq_mask(uint64_t x, uint32_t c) { return byte_mask(zero_bytes(x ^ (0x0101010101010101ULL * c))); }
// bit i set when byte i of x is a gap character (',' or a C-locale isspace:
// ' ', 9..13).  The 9..13 test adds to the low 7 bits of each byte, which
// cannot carry into the next byte.
HYD_HD uint32_t gap_mask8(uint64_t x) {
    const uint64_t ones = 0x0101010101010101ULL;
    const uint64_t lo7 = x & (ones * 0x7F);
    const uint64_t ge9 = lo7 + ones * (0x80 - 9);    // bit 7: low bits >= 9
    const uint64_t ge14 = lo7 + ones * (0x80 - 14);  // bit 7: low bits >= 14
    const uint64_t ctl = ge9 & ~ge14 & ~x;          // and the byte itself < 0x80
    return byte_mask(zero_bytes(x ^ (ones * ',')) | zero_bytes(x ^ (ones * ' ')) | ctl);
}
HYD_HD int lead_digits(uint64_t x) {  // digit bytes at the start of x (0..8)
    const uint32_t nd = ~digit_mask(x) & 0xFF;
    return nd ? __builtin_ctz(nd) : 8;
}
HYD_HD uint32_t pow10_small(int k) {  // 10^k, 0 <= k <= 8, without a table
    uint32_t p = (k & 1) ? 10u : 1u;
    p *= (k & 2) ? 100u : 1u;
    p *= (k & 4) ? 10000u : 1u;
    p *= (k & 8) ? 100000000u : 1u;
    return p;
}
// value of the first k (1..8) digit bytes of x: padded with leading '0's to
// eight digits, then pairs, quads and the whole combined by three multiplies
HYD_HD uint32_t digits_value(uint64_t x, int k) {
    if (k < 8) x = (x << (8 * (8 - k))) | (0x3030303030303030ULL >> (8 * k));
    x -= 0x3030303030303030ULL;
    x = x * 10 + (x >> 8);
    x = (((x & 0x000000FF000000FFULL) * (100 + (1000000ULL << 32))) +
         (((x >> 16) & 0x000000FF000000FFULL) * (1 + (10000ULL << 32)))) >> 32;
    return (uint32_t)x;
}

// FastSink for a run of k (<= 8) digit bytes at the start of x.
struct RunSink {
    uint64_t w = 0;
    int nsig = 0;
    int64_t qadj = 0;
    bool truncated = false;
    HYD_HD void run(uint64_t x, int k, bool after_point) {
        if (nsig == 0) {  // leading zeros only move the point
            const uint32_t nz = ~eq_mask(x, '0') & 0xFF;
            int z = nz ? __builtin_ctz(nz) : 8;
            z = z < k ? z : k;
            if (after_point) qadj -= z;
            x = z < 8 ? (x >> (8 * z)) : 0;
            k -= z;
            if (k == 0) return;
        }
        int take = FastSink::kMaxDec - nsig;
        take = take < k ? take : k;
        if (take > 0) {
            w = w * pow10_small(take) + digits_value(x, take);
            nsig += take;
            if (after_point) qadj -= take;
        }
        const int rest = k - take;
        if (rest > 0) {
            if (!after_point) qadj += rest;
            const uint32_t dropped = ((1u << k) - 1) & ~((1u << take) - 1);
            if (~eq_mask(x, '0') & dropped) truncated = true;
        }
    }
};

// rd8(q): the 8 bytes at q..q+7 (valid for q < p + kSwarSpan); at(q): one byte.
template <class Rd, class At>
HYD_HD bool scan_dec_swar(const Rd& rd8, const At& at, int64_t p, FloatTok* t) {
    int64_t i = p;
    int c = at(i);
    bool neg = false;
    if (c == '+' || c == '-') {
        neg = c == '-';
        c = at(++i);
    }
    if (c == '0' && lower(at(i + 1)) == 'x') return false;  // hexadecimal
    RunSink s;
    int64_t mant = 0;
    for (int part = 0; part < 2; ++part) {  // before and after the point
        for (;;) {
            if (i - p > kSwarSpan - 8) return false;
            const uint64_t x = rd8(i);
            const int k = lead_digits(x);
            if (k) s.run(x, k, part == 1);
            mant += k;
            i += k;
            if (k < 8) break;
        }
        if (part == 1 || at(i) != '.') break;
        ++i;
    }
    if (mant == 0) return false;  // "." , "-.", letters
    int64_t e = 0;
    if (lower(at(i)) == 'e') {
        int64_t j = i + 1;
        const int cs = at(j);
        const bool eneg = cs == '-';
        if (cs == '+' || cs == '-') ++j;
        if (j - p > kSwarSpan - 8) return false;
        const uint64_t x = rd8(j);
        const int k = lead_digits(x);
        if (k == 8) return false;  // eight or more exponent digits
        if (k) e = digits_value(x, k);
        if (eneg) e = -e;
        i = j + k;  // no digits: the marker (and sign) are taken but ignored
    }
    t->len = i - p;
    t->kind = kDec;
    t->neg = neg;
    t->w = s.w;
    t->q = e + s.qadj;
    t->truncated = s.truncated;
    return true;
}

// Bits of a scanned token when the table path decides them; false when the
// exact path must (decimal tokens only).
HYD_HD bool tok_to_bits(const FloatTok& t, uint64_t* out) {
    uint64_t mag;
    bool ok = true;
    switch (t.kind) {
        case kNan: mag = kNanBits; break;
        case kInf: mag = kInfBits; break;
        case kHexZero: mag = 0; break;
        case kHex: mag = hex_to_bits(t.w, t.q, t.truncated); break;
        default:
            if (t.w == 0) {
                mag = 0;
            } else if (!t.truncated) {
                ok = dec_to_bits(t.w, t.q, &mag);
            } else {
                // x lies strictly between w * 10^q and (w + 1) * 10^q
                uint64_t a, b;
                const bool oa = dec_to_bits(t.w, t.q, &a);
                const bool ob = dec_to_bits(t.w + 1, t.q, &b);
                ok = oa && ob && a == b;
                mag = a;
            }
    }
    *out = mag | (t.neg ? kSignBit : 0);
    return ok;
}

// ---- exact big-integer comparison (rare; a separate kernel) -----------------
constexpr int kBigLimbs = 100;  // 3200 bits: every comparison below stays under 2800
constexpr int kMaxDigits = 800; // significant digits kept; a midpoint has fewer than 770

struct Big {
    uint32_t d[kBigLimbs];
    int n;  // limbs in use; d[n-1] != 0 unless n == 0
};

HYD_HD void big_set(Big& b, uint64_t v) {
    b.n = 0;
    while (v) {
        b.d[b.n++] = (uint32_t)v;
        v >>= 32;
    }
}

HYD_HD void big_mul_add(Big& b, uint32_t m, uint32_t a) {
    uint64_t carry = a;
    for (int i = 0; i < b.n; ++i) {
        const uint64_t t = (uint64_t)b.d[i] * m + carry;
        b.d[i] = (uint32_t)t;
        carry = t >> 32;
    }
    if (carry && b.n < kBigLimbs) b.d[b.n++] = (uint32_t)carry;
}

HYD_HD void big_mul_pow5(Big& b, int64_t k) {
    const uint32_t p13 = 1220703125u;  // 5^13
    while (k >= 13) {
        big_mul_add(b, p13, 0);
        k -= 13;
    }
    uint32_t m = 1;
    while (k-- > 0) m *= 5;
    if (m != 1) big_mul_add(b, m, 0);
}

HYD_HD int64_t big_bitlen(const Big& b) {
    if (b.n == 0) return 0;
    return (int64_t)(b.n - 1) * 32 + (32 - __builtin_clz(b.d[b.n - 1]));
}

HYD_HD void big_shl(Big& b, int64_t s) {
    if (b.n == 0 || s <= 0) return;
    const int limbs = (int)(s / 32), bits = (int)(s % 32);
    int n = b.n + limbs + 1;
    if (n > kBigLimbs) n = kBigLimbs;
    for (int i = n - 1; i >= 0; --i) {
        const int src = i - limbs;
        uint32_t v = 0;
        if (src >= 0 && src < b.n) v = b.d[src] << bits;
        if (bits && src - 1 >= 0 && src - 1 < b.n) v |= b.d[src - 1] >> (32 - bits);
        b.d[i] = v;
    }
    b.n = n;
    while (b.n && b.d[b.n - 1] == 0) --b.n;
}

HYD_HD int big_cmp(const Big& a, const Big& b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}

// sign of (D + sticky * eps) * 10^Q * 2^H - C * 2^G, where D > 0, C > 0 and
// 0 < eps < 1.  L and R are work space.
HYD_HD int cmp_scaled(const Big& D, int64_t Q, int64_t H, uint64_t C, int64_t G, bool sticky, Big& L,
                      Big& R) {
    L = D;
    big_set(R, C);
    if (Q >= 0) big_mul_pow5(L, Q);
    else big_mul_pow5(R, -Q);
    const int64_t t = Q + H - G;  // compare L * 2^t with R
    const int64_t bl = big_bitlen(L) + t, br = big_bitlen(R);
    int c;
    if (bl != br) {
        c = bl < br ? -1 : 1;
    } else {
        if (t >= 0) big_shl(L, t);
        else big_shl(R, -t);
        c = big_cmp(L, R);
    }
    if (c == 0 && sticky) c = 1;
    return c;
}

// Sink keeping up to kMaxDigits significant digits in a Big (9 at a time).
struct BigSink {
    Big* D;
    int nsig = 0;
    int64_t qadj = 0;
    bool truncated = false;
    uint32_t chunk = 0;
    int chunk_len = 0;
    HYD_HD void digit(int d, bool after_point, bool /*hex*/) {
        if (nsig == 0 && d == 0) {
            if (after_point) qadj -= 1;
        } else if (nsig < kMaxDigits) {
            chunk = chunk * 10 + (uint32_t)d;
            if (++chunk_len == 9) flush();
            ++nsig;
            if (after_point) qadj -= 1;
        } else {
            if (d) truncated = true;
            if (!after_point) qadj += 1;
        }
    }
    HYD_HD void flush() {
        if (chunk_len == 0) return;
        uint32_t m = 1;
        for (int i = 0; i < chunk_len; ++i) m *= 10;
        big_mul_add(*D, m, chunk);
        chunk = 0;
        chunk_len = 0;
    }
};

HYD_HD void mid_of(uint64_t a, uint64_t* C, int64_t* G) {  // midpoint of a and its successor
    const uint64_t frac = a & ((1ULL << 52) - 1);
    const int be = (int)(a >> 52);
    const uint64_t m = be ? (frac | (1ULL << 52)) : frac;
    const int64_t e = be ? be - 1075 : -1074;
    *C = 2 * m + 1;
    *G = e - 1;
}

// Exact binary64 bits of the decimal token at p (kind kDec), given its table
// candidate `cand` (magnitude within one unit of the answer).  L, R, D are
// work space.
template <class At>
HYD_HD uint64_t dec_exact_bits(const At& at, int64_t p, uint64_t cand, Big& D, Big& L, Big& R) {
    D.n = 0;
    BigSink sink;
    sink.D = &D;
    const FloatTok t = scan_f64(at, p, sink);
    sink.flush();
    const uint64_t sign = t.neg ? kSignBit : 0;
    if (D.n == 0) return sign;
    const int64_t Q = t.q;  // x = (D + sticky) * 10^Q
    if (sink.nsig + Q > 310) return sign | kInfBits;
    if (sink.nsig + Q < -343) return sign;
    uint64_t z = cand & ~kSignBit;
    for (int it = 0; it < 8; ++it) {
        uint64_t C;
        int64_t G;
        if (z < kInfBits) {
            mid_of(z, &C, &G);
            const int c = cmp_scaled(D, Q, 0, C, G, sink.truncated, L, R);
            if (c > 0 || (c == 0 && (z & 1))) {
                ++z;
                continue;
            }
        }
        if (z > 0) {
            mid_of(z - 1, &C, &G);
            const int c = cmp_scaled(D, Q, 0, C, G, sink.truncated, L, R);
            if (c < 0 || (c == 0 && !((z - 1) & 1))) {
                --z;
                continue;
            }
        }
        break;
    }
    return sign | z;
}

// ---- binary64 -> 17 significant digits --------------------------------------
// |v| (finite, nonzero magnitude bits): N in [10^16, 10^17) and X with
// |v| ~ N * 10^(X - 16).  Returns false when the exact check must decide the
// rounding: then N holds the table's floor of |v| * 10^(16 - X) (exact, or one
// above a value less than half a unit below it) and X the scale used.
// FORCE (test builds only) hands every value to the exact check.
// The estimate x0 = floor(e2 * log10 2) of a finite nonzero |v| = m * 2^e
// (e2 = floor(log2 |v|)), from which f64_to_dec17 starts: its X is x0, x0 + 1
// (the estimate one low) or x0 + 2 (17 nines rounded up to 10^17).
HYD_HD int dec17_x0(uint64_t bits) {
    const uint64_t frac = bits & ((1ULL << 52) - 1);
    const int be = (int)((bits >> 52) & 0x7FF);
    const uint64_t m = be ? (frac | (1ULL << 52)) : frac;
    const int e = be ? be - 1075 : -1074;
    const int e2 = e + (63 - clz64(m));
    return (e2 * 78913) >> 18;
}

template <bool FORCE = false>
HYD_HD bool f64_to_dec17(uint64_t bits, uint64_t* N, int* X) {
    const uint64_t frac = bits & ((1ULL << 52) - 1);
    const int be = (int)((bits >> 52) & 0x7FF);
    const uint64_t m = be ? (frac | (1ULL << 52)) : frac;
    const int e = be ? be - 1075 : -1074;
    const int lz = clz64(m);
    int x0 = dec17_x0(bits);  // floor(e2 * log10(2)): x0 <= X <= x0 + 1 before rounding
    const uint64_t mn = m << lz;
    for (int pass = 0;; ++pass) {
        const int s = 16 - x0;
        const uint64_t* row = pow5_row(s);
        const U192 P = mul_64x128(mn, row[0], row[1]);
        const int k = -((int)(int64_t)row[2] + e + s - lz);  // |v| * 10^s ~ P / 2^k
        const uint64_t nf = shr_192(P, k);
        if (nf >= kTen17 && pass == 0) {
            ++x0;
            continue;
        }
        const bool rb = bit_at(P, k - 1);
        bool up;
        if (FORCE) {
            *N = nf;
            *X = x0;
            return false;
        }
        if (s >= 0 && s <= kExactMax) {
            up = rb && (!bits_zero(P, 0, k - 1) || (nf & 1));
        } else {
            if (rb && bits_zero(P, 64, k - 1) && P.w0 < mn) {
                *N = nf;
                *X = x0;
                return false;
            }
            up = rb;
        }
        uint64_t n = nf + (up ? 1 : 0);
        if (n == kTen17) {
            n = kTen16;
            ++x0;
        }
        *N = n;
        *X = x0;
        return true;
    }
}

// Settles an undecided f64_to_dec17: nf = floor(|v| * 10^(16 - x)).
HYD_HD void dec17_exact(uint64_t bits, uint64_t nf, int x, uint64_t* N, int* X, Big& D, Big& L, Big& R) {
    const uint64_t frac = bits & ((1ULL << 52) - 1);
    const int be = (int)((bits >> 52) & 0x7FF);
    const uint64_t m = be ? (frac | (1ULL << 52)) : frac;
    const int e = be ? be - 1075 : -1074;
    big_set(D, m);
    // |v| * 10^s against nf + 1/2
    const int c = cmp_scaled(D, 16 - x, e, 2 * nf + 1, -1, false, L, R);
    uint64_t n = nf + ((c > 0 || (c == 0 && (nf & 1))) ? 1 : 0);
    if (n == kTen17) {
        n = kTen16;
        ++x;
    }
    *N = n;
    *X = x;
}

// ---- text writers ---------------------------------------------------------
HYD_HD int dec_digits(uint64_t v) {  // decimal digits of v >= 1
    int n = 1;
    while (v >= 10) {
        v /= 10;
        ++n;
    }
    return n;
}

HYD_HD int i64_len(int64_t v) {
    const uint64_t a = v < 0 ? 0 - (uint64_t)v : (uint64_t)v;
    return (v < 0) + (a ? dec_digits(a) : 1);
}

// 8 ASCII digits of v < 10^8, the most significant in the lowest byte: the two
// 4-digit halves, then their 2-digit halves, then single digits, each split
// by a multiply-shift division in parallel 32- / 16-bit lanes (exact for the
// ranges involved; checked for every v < 10^8).
HYD_HD uint64_t ascii8(uint32_t v) {
    uint64_t x = (uint64_t)(v / 10000) | ((uint64_t)(v % 10000) << 32);
    const uint64_t y = ((x * 10486) >> 20) & 0x0000007F0000007FULL;  // x / 100 per lane (x < 10^4)
    x = y | ((x - y * 100) << 16);
    const uint64_t z = ((x * 103) >> 10) & 0x000F000F000F000FULL;    // x / 10 per lane (x < 100)
    x = z | ((x - z * 10) << 8);
    return x | 0x3030303030303030ULL;
}

// Writes %lld of v at pos (Pos: the caller's position type - 32-bit offsets
// into a workgroup's staging buffer on the device).  The 19-digit decimal
// string of |v| is 3 leading digits and two 8-digit ASCII words (ascii8); the
// last len of its digits are written, each at a constant offset from one base.
template <class Put, class Pos>
HYD_HD int i64_write(int64_t v, const Put& put, Pos pos) {
    const uint64_t a = v < 0 ? 0 - (uint64_t)v : (uint64_t)v;
    const int nd = a ? dec_digits(a) : 1;
    const int len = (v < 0) + nd;
    if (v < 0) put(pos, '-');
    const uint32_t top = (uint32_t)(a / 10000000000000000ULL);  // < 923
    const uint64_t rem = a - (uint64_t)top * 10000000000000000ULL;
    const uint32_t mid = (uint32_t)(rem / 100000000ULL), low = (uint32_t)(rem - (uint64_t)mid * 100000000ULL);
    const uint64_t am = ascii8(mid), al = ascii8(low);
    const Pos base = pos + (Pos)(len - 19);  // digit k of the 19 goes to base + k
#pragma unroll
    for (int k = 0; k < 19; ++k) {
        if (k < 19 - nd) continue;
        char ch;
        if (k == 0) ch = (char)('0' + top / 100);
        else if (k == 1) ch = (char)('0' + (top / 10) % 10);
        else if (k == 2) ch = (char)('0' + top % 10);
        else if (k < 11) ch = (char)((am >> (8 * (k - 3))) & 0xFF);
        else ch = (char)((al >> (8 * (k - 11))) & 0xFF);
        put(base + (Pos)k, ch);
    }
    return len;
}


// A finite nonzero value's 17 digits N (from f64_to_dec17) and exponent X:
// digit 0, then digits 1..8 and 9..16 as ASCII words; nd = digits left once
// the trailing zeros are dropped.
struct G17 {
    uint64_t a0, a1;  // ASCII digits 1..8 and 9..16 (digit 1 in the low byte)
    int d0;           // leading digit, 1..9
    int nd;
    int X;
    bool neg;
};

HYD_HD int trailing_zero_digits(uint64_t a) {  // '0' bytes at the top end of an ASCII word
    const uint32_t other = ~eq_mask(a, '0') & 0xFF;
    return other ? __builtin_clz(other) - 24 : 8;
}

HYD_HD G17 g17_make(bool neg, uint64_t N, int X) {
    G17 g;
    g.neg = neg;
    g.X = X;
    g.d0 = (int)(N / kTen16);
    const uint64_t r = N - (uint64_t)g.d0 * kTen16;
    const uint32_t hi = (uint32_t)(r / 100000000ULL), lo = (uint32_t)(r - (uint64_t)hi * 100000000ULL);
    g.a0 = ascii8(hi);
    g.a1 = ascii8(lo);
    const int tz = lo ? trailing_zero_digits(g.a1) : (hi ? 8 + trailing_zero_digits(g.a0) : 16);
    g.nd = 17 - tz;
    return g;
}

HYD_HD bool g17_exp_form(int X) { return X < -4 || X >= 17; }

HYD_HD int g17_len(const G17& g) {
    const int s = g.neg ? 1 : 0;
    if (g17_exp_form(g.X)) {
        const int ax = g.X < 0 ? -g.X : g.X;
        return s + 1 + (g.nd > 1 ? g.nd : 0) + 2 + (ax >= 100 ? 3 : 2);
    }
    if (g.X >= 0) return s + (g.nd > g.X + 1 ? g.nd + 1 : g.X + 1);
    return s + 2 + (-g.X - 1) + g.nd;
}

HYD_HD char g17_digit(const G17& g, int j) {  // j is a constant after unrolling
    if (j == 0) return (char)('0' + g.d0);
    return (char)((j <= 8 ? g.a0 >> (8 * (j - 1)) : g.a1 >> (8 * (j - 9))) & 0xFF);
}

template <class Put, class Pos>
HYD_HD int g17_write(const G17& g, const Put& put, Pos pos) {
    const int len = g17_len(g);
    const int s = g.neg ? 1 : 0;
    if (g.neg) put(pos, '-');
    const bool ef = g17_exp_form(g.X);
    // digit j goes to pos + s + j + gap(j); digits written: nd, or the whole
    // integer part of the %f form when it is longer (its tail digits are 0s)
    int ndw, lead, dot_after;  // dot after digit dot_after (-1: none); lead: chars before digit 0
    if (ef) {
        ndw = g.nd;
        lead = 0;
        dot_after = g.nd > 1 ? 0 : -1;
    } else if (g.X >= 0) {
        ndw = g.nd > g.X + 1 ? g.nd : g.X + 1;
        lead = 0;
        dot_after = g.nd > g.X + 1 ? g.X : -1;
    } else {
        ndw = g.nd;
        lead = 1 - g.X;  // "0." and -X-1 zeros
        put(pos + s, '0');
        put(pos + s + 1, '.');
        for (int i = 0; i < -g.X - 1; ++i) put(pos + s + 2 + i, '0');
        dot_after = -1;
    }
    const Pos d = pos + (Pos)(s + lead);
#pragma unroll
    for (int j = 0; j < 17; ++j)
        if (j < ndw) put(d + j + (dot_after >= 0 && j > dot_after ? 1 : 0), g17_digit(g, j));
    if (dot_after >= 0) put(d + dot_after + 1, '.');
    if (ef) {
        Pos q = d + (Pos)(ndw + (dot_after >= 0 ? 1 : 0));
        put(q++, 'e');
        put(q++, g.X < 0 ? '-' : '+');
        const int ax = g.X < 0 ? -g.X : g.X;
        if (ax >= 100) put(q++, (char)('0' + ax / 100));
        put(q++, (char)('0' + (ax / 10) % 10));
        put(q, (char)('0' + ax % 10));
    }
    return len;
}

// Specials and zero: length of "%.17g" text, or 0 for a finite nonzero value.
HYD_HD int special_len(uint64_t bits) {
    const bool neg = bits >> 63;
    const uint64_t mag = bits & ~kSignBit;
    if (mag == 0) return 1 + neg;
    if (mag >= kInfBits) return 3 + neg;
    return 0;
}

template <class Put, class Pos>
HYD_HD void special_write(uint64_t bits, const Put& put, Pos pos) {
    const bool neg = bits >> 63;
    const uint64_t mag = bits & ~kSignBit;
    if (neg) put(pos++, '-');
    if (mag == 0) {
        put(pos, '0');
    } else if (mag == kInfBits) {
        put(pos, 'i');
        put(pos + 1, 'n');
        put(pos + 2, 'f');
    } else {
        put(pos, 'n');
        put(pos + 1, 'a');
        put(pos + 2, 'n');
    }
}

}  // namespace csv
}  // namespace hyd
