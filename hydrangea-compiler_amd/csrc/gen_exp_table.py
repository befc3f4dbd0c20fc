"""Generate the 2^(k/128) table used by hyd_exp.h from exact decimal arithmetic.

Entry k is a pair of 64-bit words:
  word 0 = bits of T[k], the relative tail (2^(k/128) - H[k]) / H[k] rounded
           to the nearest double;
  word 1 = bits of H[k] (2^(k/128) rounded to nearest) minus (k << 45), so that
           adding (ki << 45) for an integer ki with ki % 128 == k rebuilds the
           exponent of 2^(ki/128).
The table is fixed by the mathematics: rerunning this script reproduces the
committed hyd_exp_table.inc bit for bit.
"""
import os
import struct
from decimal import Decimal, getcontext

N = 128


def _bits(d: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def table():
    getcontext().prec = 80
    ln2 = Decimal(2).ln()
    rows = []
    for k in range(N):
        v = (ln2 * k / N).exp()
        h = float(v)  # Decimal -> float conversion is correctly rounded
        t = float((v - Decimal(h)) / Decimal(h))
        rows.append((_bits(t), (_bits(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    return rows


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "include", "hyd_exp_table.inc")
    with open(out, "w") as f:
        f.write("/* generated by gen_exp_table.py: 2^(k/128) as (tail, scale-bits) pairs */\n")
        for t, h in table():
            f.write("0x%016xULL, 0x%016xULL,\n" % (t, h))


if __name__ == "__main__":
    main()
