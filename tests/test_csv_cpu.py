"""CSV text <-> arrays on the CPU (no GPU):

* the oracle (the reference's read/write loops of
  runtime/hydrangea_runtime.h:549-594 and runtime/hyd_write_csv.c:5-59 over
  glibc's fscanf / fprintf) against the reference's own CSV fixtures and known
  answers, and against Python's float parsing and '%.17g' formatting, an
  independent correctly rounded implementation (David Gay's dtoa);
* the conversions the GPU kernels use (csv_num.h, compiled for the host into
  libhydhip's test hooks) against the oracle, on fuzzed and adversarial
  numbers, with and without forcing the exact big-integer path.
"""
import os
import random
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "csv")


def gold(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def host():
    import hydhip  # the library loads without a GPU; only the host hooks run

    hydhip.lib()
    return hydhip


# ---- oracle pins --------------------------------------------------------------
# examples/*.csv of the reference, read as its example programs read them
FIXTURES = [
    ("data.csv", 6, False, [1, 2, 3, 4, 5, 6]),             # examples/read_array.hyd
    ("dotA.csv", 3, False, [1, 2, 3]),                      # examples/dot_csv.hyd
    ("dotB.csv", 3, False, [4, 5, 6]),
    ("matA.csv", 6, True, [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]),
    ("matB.csv", 6, True, [7.0, 8.0, 9.0, 10.0, 11.0, 12.0]),
    ("matA_3x3.csv", 9, True, [float(k) for k in range(1, 10)]),
    ("matB_3x3.csv", 9, True, [float(k) for k in range(9, 0, -1)]),
    ("matA_4x4.csv", 16, True, [float(k) for k in range(1, 17)]),
    ("matB_4x4.csv", 16, True, [1.0 if i % 5 == 0 else 0.0 for i in range(16)]),
]


@pytest.mark.parametrize("name,count,is_float,expected", FIXTURES)
def test_oracle_reads_reference_fixtures(orc, name, count, is_float, expected):
    vals, kind, _ = orc.csv_read(gold(name), count, is_float)
    assert kind == orc.CSV_OK
    assert vals.tolist() == expected


def test_oracle_reference_known_answers(orc):
    # test/Language/Hydrangea/InterpreterSpec.hs:402-440
    assert orc.csv_read(b"10,20,30", 3, False)[0].tolist() == [10, 20, 30]
    assert orc.csv_read(b"1,2,3,4,5,6", 6, False)[0].tolist() == [1, 2, 3, 4, 5, 6]
    assert int(orc.csv_read(b"1,2,3,4", 4, False)[0].sum()) == 10
    assert orc.csv_read(b"1,two,3", 3, False)[1:] == (orc.CSV_PARSE_ERROR, 1)
    assert orc.csv_read(b"1.0,oops,3.0", 3, True)[1:] == (orc.CSV_PARSE_ERROR, 1)
    # test/Language/Hydrangea/EndToEndSpec.hs:241-259: write_array_float of a
    # 2 x 3 array keeps the rows
    vals = np.arange(6, dtype=np.float64) + 0.5
    assert orc.csv_write(vals, 2, 3, True) == b"0.5,1.5,2.5\n3.5,4.5,5.5\n"


# The reference's loop (runtime/hydrangea_runtime.h:572-594) read by hand:
# between elements it skips ',' ' ' '\n' '\r', then fscanf skips isspace.
LOOP_CASES = [
    (b"1,2,,3", 3, False, (0, -1), [1, 2, 3]),
    (b" \n\t1,2", 2, False, (0, -1), [1, 2]),
    (b"1-2+3", 3, False, (0, -1), [1, -2, 3]),
    (b"1\r\n2\r\n", 2, False, (0, -1), [1, 2]),
    (b"1,\t2", 2, False, (0, -1), [1, 2]),        # ',' then isspace
    (b"1\t,2", 2, False, (1, 1), None),           # isspace then ',': fscanf meets ','
    (b",1", 1, False, (1, 0), None),              # only isspace before the first
    (b"1,2,", 3, False, (2, 2), None),            # separator loop hits the end
    (b"1,2,\t", 3, False, (1, 2), None),          # fscanf hits the end
    (b"", 1, False, (1, 0), None),
    (b"", 0, False, (0, -1), []),
    (b"1,2,3,junk", 3, False, (0, -1), [1, 2, 3]),  # text after the last element is not read
    (b"1e,2", 2, True, (0, -1), [1.0, 2.0]),      # glibc: "1e" reads 1
    (b"1.5.5", 2, True, (0, -1), [1.5, 0.5]),
    (b"nan(1),2", 2, True, (1, 1), None),         # the "(1)" is not read
    (b"0x1p3 -inf", 2, True, (0, -1), [8.0, -np.inf]),
]


@pytest.mark.parametrize("text,count,is_float,status,expected", LOOP_CASES)
def test_oracle_loop_rules(orc, text, count, is_float, status, expected):
    vals, kind, elem = orc.csv_read(text, count, is_float)
    assert (kind, elem) == status
    if expected is not None:
        assert vals.tolist() == expected


def random_doubles(rng, n):
    b = rng.integers(0, 2**63, n, dtype=np.int64).view(np.uint64)
    b |= rng.integers(0, 2, n).astype(np.uint64) << np.uint64(63)
    v = b.view(np.float64)
    return v[np.isfinite(v)]


def near_midpoints(rng, n):
    """Decimal strings at, or one unit in the last place beside, the exact
    midpoint of two neighbouring doubles, with 17 to 800 significant digits."""
    out = []
    R = random.Random(int(rng.integers(1 << 30)))
    while len(out) < n:
        x = float(rng.integers(1, 2**53)) * 2.0 ** int(rng.integers(-1126, 970))
        nx = np.nextafter(x, np.inf)
        if x == 0 or not np.isfinite(nx):
            continue
        mid = (Fraction(x) + Fraction(nx)) / 2
        d = R.choice([17, 18, 19, 20, 25, 40, 120, 400, 800, 805])
        e10 = int((mid.numerator.bit_length() - mid.denominator.bit_length()) * 0.30103)
        m = mid * Fraction(10) ** (d - 1 - e10)
        k = m.numerator // m.denominator + R.choice([-1, 0, 0, 1])
        out.append(("%de%d" % (k, -(d - 1 - e10))).encode())
    return out


def test_oracle_formats_like_python(orc):
    rng = np.random.default_rng(11)
    v = np.concatenate([random_doubles(rng, 50_000), [0.0, -0.0, 5e-324, 2.0**-25, 1e16, 1e17, 1e-5, 1e-4,
                                                    1.7976931348623157e308]])
    got = orc.format_values(v, True)
    assert got == [("%.17g" % x).encode() for x in v]


def test_oracle_parses_like_python(orc):
    rng = np.random.default_rng(12)
    R = random.Random(12)
    strs = near_midpoints(rng, 4000)
    for _ in range(20_000):
        digs = "".join(R.choice("0123456789") for _ in range(R.randint(1, 30)))
        pt = R.randint(0, len(digs))
        s = digs[:pt] + "." + digs[pt:] + R.choice(["", "e%d" % R.randint(-340, 310)])
        strs.append((R.choice(["", "-", "+"]) + s).encode())
    r, vals, consumed = orc.scan_tokens(strs, True)
    assert (r == 1).all() and (consumed == [len(s) for s in strs]).all()
    assert (bits(vals) == bits([float(s) for s in strs])).all()
    hexes = ["0x%x.%xp%d" % (R.getrandbits(60), R.getrandbits(40), R.randint(-1130, 960)) for _ in range(5000)]
    r, vals, _ = orc.scan_tokens([h.encode() for h in hexes], True)
    assert (bits(vals) == bits([float.fromhex(h) for h in hexes])).all()


# ---- the device conversions, built for the host, against the oracle ----------------
def check_scan(host, orc, strs, is_float, force_exact=False, swar=False):
    r, v, c, _ = host.host_csv_scan(strs, is_float, force_exact, swar)
    r2, v2, c2 = orc.scan_tokens(strs, is_float)
    assert (r == r2).all()
    ok = r == 1
    assert (c[ok] == c2[ok]).all()
    if is_float:
        assert (bits(v)[ok] == bits(v2)[ok]).all()
    else:
        assert (v[ok] == v2[ok]).all()


@pytest.mark.parametrize("force_exact", [False, True])
def test_device_formatter_matches_glibc(host, orc, force_exact):
    rng = np.random.default_rng(21)
    v = np.concatenate([random_doubles(rng, 100_000),
                        [2.0**e for e in range(-1074, 1024)],
                        [float("%de%d" % (d, e)) for e in range(-323, 308, 7) for d in (1, 5, 9999999999999999)],
                        [0.0, -0.0, np.inf, -np.inf]])
    got, exact = host.host_csv_format(v, True, force_exact=force_exact)
    assert got == orc.format_values(v, True)
    if force_exact:
        assert exact == np.count_nonzero(np.isfinite(v) & (v != 0))
    # every digit count: 10^k - 1, 10^k, 10^k + 1 and their negatives (the
    # writer splits |v| into 3 + 8 + 8 digits), small values, the extremes
    pw = [10**k + d for k in range(19) for d in (-1, 0, 1)]
    edges = np.array([x for p in pw for x in (p, -p) if -2**63 <= x < 2**63], np.int64)
    i = np.concatenate([rng.integers(-2**63, 2**63 - 1, 20_000, dtype=np.int64),
                        rng.integers(-1000, 1000, 2000, dtype=np.int64), edges,
                        np.array([0, -1, 2**63 - 1, -2**63, -2**63 + 1], np.int64)])
    assert host.host_csv_format(i, False)[0] == orc.format_values(i, False)


# swar: the 8-characters-at-a-time path the value kernel tries first
@pytest.mark.parametrize("force_exact,swar", [(False, False), (True, False), (False, True), (True, True)])
def test_device_scanner_matches_glibc(host, orc, force_exact, swar):
    rng = np.random.default_rng(22)
    R = random.Random(22)
    check_scan(host, orc, orc.format_values(random_doubles(rng, 50_000), True), True, force_exact, swar)
    check_scan(host, orc, near_midpoints(rng, 6000), True, force_exact, swar)
    strs = []
    for _ in range(30_000):
        digs = "".join(R.choice("0123456789") for _ in range(R.randint(1, 30)))
        if R.random() < 0.3:
            digs = "0" * R.randint(1, 40) + digs
        pt = R.randint(0, len(digs))
        s = digs[:pt] + ("." if R.random() < 0.7 else "") + digs[pt:]
        if R.random() < 0.6:
            s += R.choice("eE") + R.choice(["", "+", "-"]) + str(R.randint(0, 400))
        strs.append((R.choice(["", "", "+", "-"]) + s).encode())
    for _ in range(15_000):
        digs = "".join(R.choice("0123456789abcdefABCDEF") for _ in range(R.randint(0, 30)))
        pt = R.randint(0, len(digs))
        s = "0" + R.choice("xX") + digs[:pt] + ("." if R.random() < 0.5 else "") + digs[pt:]
        if R.random() < 0.6:
            s += R.choice("pP") + R.choice(["", "+", "-"]) + str(R.randint(0, 1200))
        strs.append((R.choice(["", "-"]) + s).encode())
    check_scan(host, orc, strs, True, force_exact, swar)


def test_device_scanner_matches_glibc_on_malformed_text(host, orc):
    R = random.Random(23)
    alpha = list("0123456789.eE+-xXpPaAfFiInNtTyY,( \t") + ["0x", "inf", "nan", "infinity", "e+", "1."]
    strs = ["".join(R.choice(alpha) for _ in range(R.randint(1, 12))).encode() for _ in range(60_000)]
    strs += [b"1e,", b"1e+,", b"0x", b"0x.", b"0x.p1", b"0xp3", b"0x1p", b"infinit,", b"infx", b"nan(1)", b".",
             b"-.", b"-", b"+,", b"1..", b"1ee5", b"--1", b" ", b""]
    check_scan(host, orc, strs, True)
    check_scan(host, orc, strs, True, swar=True)
    ints = ["".join(R.choice("0123456789+- x.,\t") for _ in range(R.randint(1, 25))).encode()
            for _ in range(30_000)]
    ints += [b"99999999999999999999", b"-99999999999999999999", b"9223372036854775808",
             b"-9223372036854775808", b"0x10", b"+5", b"-", b"00012", b"-0"]
    check_scan(host, orc, ints, False)


def test_gap_classifier_all_bytes(host):
    # the 8-byte gap classifier of the word kernels against the definition
    # (',' or C-locale isspace), every byte value at every position among
    # random neighbours
    import ctypes
    rng = np.random.default_rng(31)
    buf = rng.integers(0, 256, 8 * 256 * 8 * 16, dtype=np.uint8)
    for v in range(256):
        for pos in range(8):
            for rep in range(16):
                buf[((v * 8 + pos) * 16 + rep) * 8 + pos] = v
    out = np.zeros_like(buf)
    host.lib().hydhip_host_csv_gap_flags(buf.ctypes.data, buf.size, out.ctypes.data)
    want = np.isin(buf, [ord(","), ord(" "), 9, 10, 11, 12, 13]).astype(np.uint8)
    assert (out == want).all()
