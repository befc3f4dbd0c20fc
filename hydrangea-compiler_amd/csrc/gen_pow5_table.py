"""Generate hyd_pow5_table.inc: 128-bit scaled powers of five for the decimal
<-> binary64 conversions of the CSV kernels (csv_num.h).

For every q in [QMIN, QMAX] the table holds F(q) = 5^q / 2^f(q), scaled so that
2^127 <= F(q) < 2^128, as two 64-bit words (high, low), plus f(q):

  q >= 0:  f = bitlen(5^q) - 128   (F exact while 5^q < 2^128, i.e. q <= 55)
  q <  0:  f = -bitlen(5^-q) - 127 (F = 2^(bitlen+127) / 5^-q, never exact)

Inexact entries are rounded UP (ceiling).  With w' * ceil(F) >= w' * F the
product over-estimates by less than w' < 2^64, so a binary64 rounding midpoint
can only hide in [P - 2^64, P] of the 192-bit product P: the kernels test for
exactly that window and send those rare values to an exact big-integer path.
Decimals that are exactly representable (0.5, 1.25, ...) therefore never look
ambiguous.  This is the table shape of the Eisel-Lemire method (D. Lemire,
"Number parsing at a gigabyte per second", Software: Practice and Experience
51(8), 2021), with the rounding direction and the ambiguity window above.

The script also checks the floor(log10(2^e)) estimate (e * 78913) >> 18 used by
the formatter over every binary64 exponent.  Rerunning it reproduces the
committed file bit for bit.
"""
import os

QMIN = -342   # 10^-343 * (2^64 - 1) < 2^-1075: any smaller decimal rounds to 0
QMAX = 341    # parsing needs q <= 308; formatting scales by 10^(16 - X) up to 10^340
EXACT_MAX = 55


def entry(q):
    if q >= 0:
        p = 5 ** q
        f = p.bit_length() - 128
        if f <= 0:
            return p << -f, f, True
        hi = p >> f
        exact = (p & ((1 << f) - 1)) == 0
        return (hi if exact else hi + 1), f, exact
    p = 5 ** -q
    b = p.bit_length()
    f = -b - 127
    num = 1 << (b + 127)
    return -(-num // p), f, False


def check_log10():
    for e in range(-1074 - 64, 1024 + 1):
        est = (e * 78913) >> 18
        # exact floor(log10(2^e))
        if e >= 0:
            exact = len(str(2 ** e)) - 1
        else:
            exact = -len(str(2 ** -e))
        assert est == exact, (e, est, exact)


def main():
    check_log10()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hyd_pow5_table.inc")
    lines = ["/* generated by gen_pow5_table.py: rows {high, low, f} for q = %d .. %d, where"
             " high:low = ceil(5^q / 2^f) and 2^127 <= high:low < 2^128 (exact for 0 <= q <= %d) */"
             % (QMIN, QMAX, EXACT_MAX)]
    for q in range(QMIN, QMAX + 1):
        v, f, exact = entry(q)
        assert (1 << 127) <= v < (1 << 128), q
        assert exact == (0 <= q <= EXACT_MAX), q
        lines.append("{0x%016xULL, 0x%016xULL, (uint64_t)(%dLL)},  /* q=%d */"
                     % (v >> 64, v & ((1 << 64) - 1), f, q))
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
