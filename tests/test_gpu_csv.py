"""GPU parity of the CSV path (read_array / read_array_float / write_array /
write_array_float: runtime/hydrangea_runtime.h:549-594,
runtime/hyd_write_csv.c:5-59) against the oracle, the reference's loops over
glibc: values bit-exact, written text byte-exact, and the same failing
element and message kind for every text the reference rejects."""
import ctypes
import os
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "csv")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def same(a, b, is_float):
    return (bits(a) == bits(b)).all() if is_float else (np.asarray(a) == np.asarray(b)).all()


def check_parse(hyd, orc, text, count, is_float):
    out, st = hyd.csv_parse(text, count, is_float)
    ref, kind, elem = orc.csv_read(text, count, is_float)
    assert (st.error_kind, st.error_element) == (kind, elem), (text[:200], count)
    if kind == 0 and count:
        assert same(out.numpy(), ref, is_float)
    out.free()
    return st


def csv_text(orc, vals, rows, row_len, is_float):
    return orc.csv_write(vals, rows, row_len, is_float)


def random_doubles(rng, n):
    b = rng.integers(0, 2**64, n, dtype=np.uint64)
    return b.view(np.float64)


@pytest.mark.parametrize("name,count,is_float", [
    ("data.csv", 6, False), ("dotA.csv", 3, False), ("dotB.csv", 3, False), ("matA.csv", 6, True),
    ("matB.csv", 6, True), ("matA_3x3.csv", 9, True), ("matB_3x3.csv", 9, True), ("matA_4x4.csv", 16, True),
    ("matB_4x4.csv", 16, True)])
def test_reference_fixtures(hyd, orc, name, count, is_float):
    with open(os.path.join(GOLD, name), "rb") as f:
        check_parse(hyd, orc, f.read(), count, is_float)


@pytest.mark.parametrize("n", [1, 7, 1000, 100_003, 2_000_000])
def test_parse_written_text(hyd, orc, n):
    print("parse n=%d" % n, flush=True)
    rng = np.random.default_rng(n)
    v = random_doubles(rng, n)
    rows = 1 if n < 1000 else 1000
    text = csv_text(orc, v, rows, n // rows, True) if n % rows == 0 else csv_text(orc, v, 1, n, True)
    check_parse(hyd, orc, text, n, True)
    i = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    check_parse(hyd, orc, csv_text(orc, i, 1, n, False), n, False)


def test_parse_separators_and_spacing(hyd, orc):
    rng = np.random.default_rng(5)
    R = random.Random(5)
    v = np.round(rng.standard_normal(50_000) * 10.0 ** rng.integers(-30, 30, 50_000), 6)
    toks = [("%.17g" % x) if R.random() < 0.5 else repr(float(x)) for x in v]
    gaps = [",", ", ", " ", "\n", "\r\n", ",\n", " ,  ", ",,", "\n\n\n", ", \t", "\t", "  \v "]
    text = "".join(t + R.choice(gaps) for t in toks).encode()
    check_parse(hyd, orc, text, len(toks), True)
    check_parse(hyd, orc, text, len(toks) // 2, True)
    ints = rng.integers(-10**12, 10**12, 50_000)
    text = "".join(str(x) + R.choice(gaps) for x in ints).encode()
    check_parse(hyd, orc, text, ints.size, False)


BAD = [
    (b"1,two,3", 3, False), (b"1.0,oops,3.0", 3, True),  # InterpreterSpec.hs:434-440
    (b"1\t,2", 2, False), (b",1", 1, False), (b"1,2,", 3, False), (b"1,2,\t", 3, True), (b"", 1, True),
    (b"   ", 1, False), (b"1,2,3,junk", 3, False), (b"1,2,3,junk", 4, False), (b"nan(1),2", 2, True),
    (b"1e,2", 2, True), (b"1ee5,2", 2, True), (b"0x,1", 2, True), (b"0x.,1", 2, True), (b"--1", 1, False),
    (b"1-2-3", 3, False), (b"1+-2", 2, False), (b"infinit,", 1, True), (b"1,2\n,\t3", 3, True),
    (b"1 2\t\t,3", 3, False), (b"\x00", 1, False), (b"5", 0, True),
]


@pytest.mark.parametrize("text,count,is_float", BAD)
def test_rejected_and_odd_texts(hyd, orc, text, count, is_float):
    check_parse(hyd, orc, text, count, is_float)


def test_injected_errors(hyd, orc):
    # a valid text with one bad character at a random place: same failing
    # element as the reference's loop (or none, past `count` or harmless)
    rng = np.random.default_rng(9)
    R = random.Random(9)
    v = random_doubles(rng, 20_000)
    v[~np.isfinite(v)] = 1.0
    base = csv_text(orc, v, 1, v.size, True)
    for trial in range(40):
        pos = R.randrange(len(base))
        bad = R.choice([b"x", b"\t,", b",", b"--", b".", b"e", b" ", b"\x00", b"(", b"0x"])
        text = base[:pos] + bad + base[pos:]
        print("inject %r at %d" % (bad, pos), flush=True)
        check_parse(hyd, orc, text, v.size, True)
        check_parse(hyd, orc, text, v.size + 1, True)


def test_tile_boundaries(hyd, orc):
    # numbers and gaps longer than a workgroup's halo, straddling tiles
    R = random.Random(3)
    toks = []
    for k in range(3000):
        kind = R.random()
        if kind < 0.2:
            toks.append("".join(R.choice("0123456789") for _ in range(R.randint(300, 1200))) + "e-%d" % R.randint(0, 900))
        elif kind < 0.4:
            toks.append("0." + "0" * R.randint(400, 900) + "123e%d" % R.randint(300, 700))
        else:
            toks.append("%.17g" % R.uniform(-1e6, 1e6))
    gaps = lambda: R.choice([",", " " * R.randint(1, 2000), "\n" * R.randint(1, 700) + "\t" * R.randint(0, 3)])
    text = ("".join(t + gaps() for t in toks)).encode()
    st = check_parse(hyd, orc, text, len(toks), True)
    assert st.exact_path >= 0
    text = (" " * 8190 + ",".join(toks[:50])).encode()  # first number starts at a tile's end
    check_parse(hyd, orc, text, 50, True)
    text = ("\t" * 9000 + "," + ",".join(toks[:5])).encode()  # ',' after tabs, far back
    check_parse(hyd, orc, text, 5, True)


def test_exact_path_on_gpu(hyd, orc):
    from fractions import Fraction
    rng = np.random.default_rng(4)
    R = random.Random(4)
    toks = []
    while len(toks) < 3000:
        x = float(rng.integers(1, 2**53)) * 2.0 ** int(rng.integers(-1126, 970))
        nx = np.nextafter(x, np.inf)
        if x == 0 or not np.isfinite(nx):
            continue
        mid = (Fraction(x) + Fraction(nx)) / 2
        d = R.choice([17, 19, 20, 40, 400, 800, 805])
        e10 = int((mid.numerator.bit_length() - mid.denominator.bit_length()) * 0.30103)
        m = mid * Fraction(10) ** (d - 1 - e10)
        k = m.numerator // m.denominator + R.choice([-1, 0, 0, 1])
        toks.append("%de%d" % (k, -(d - 1 - e10)))
    text = ",".join(toks).encode()
    st = check_parse(hyd, orc, text, len(toks), True)
    assert st.exact_path > 1000


def test_unaligned_text(hyd, orc):
    rng = np.random.default_rng(6)
    v = random_doubles(rng, 30_000)
    text = csv_text(orc, v, 1, v.size, True)
    dev = hyd.DeviceArray.from_numpy(np.frombuffer(b"#" + text, np.uint8))
    out = hyd.DeviceArray((v.size,), np.float64)
    st = hyd.CsvStatus()
    hyd._check(hyd.lib().hydhip_csv_parse_f64(dev.ptr + 1, len(text), v.size, out.ptr, ctypes.byref(st), None))
    assert st.error_kind == 0 and same(out.numpy(), orc.csv_read(text, v.size, True)[0], True)


@pytest.mark.parametrize("n,rows", [(1, 1), (6, 2), (1000, 1), (100_000, 100), (1_000_003, 1), (3_000_000, 1500)])
def test_format_matches_reference_writer(hyd, orc, n, rows):
    print("format n=%d rows=%d" % (n, rows), flush=True)
    rng = np.random.default_rng(n + rows)
    v = random_doubles(rng, n)
    v[: min(n, 8)] = [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1e17, 0.1][: min(n, 8)]
    d = hyd.DeviceArray.from_numpy(v)
    assert hyd.csv_format(d, n // rows) == orc.csv_write(v, rows, n // rows, True)
    i = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    di = hyd.DeviceArray.from_numpy(i)
    assert hyd.csv_format(di, n // rows) == orc.csv_write(i, rows, n // rows, False)


def test_format_exact_kernel_on_gpu(hyd, orc):
    # every value through the exact kernel (sparse digit table, block length
    # correction): same bytes as the reference writer
    rng = np.random.default_rng(77)
    v = random_doubles(rng, 300_000)
    v[:6] = [0.0, -0.0, np.inf, 5e-324, 2.0**-25, 1e23]
    d = hyd.DeviceArray.from_numpy(v)
    hyd.lib().hydhip_csv_test_force_exact(1)
    try:
        got, exact = hyd.csv_format(d, 1000, with_exact=True)
    finally:
        hyd.lib().hydhip_csv_test_force_exact(0)
    assert exact == np.count_nonzero(np.isfinite(v) & (v != 0))
    assert got == orc.csv_write(v, 300, 1000, True)
    assert hyd.csv_format(d, 1000) == got


def test_format_known_answer(hyd):
    # test/Language/Hydrangea/EndToEndSpec.hs:241-259
    d = hyd.DeviceArray.from_numpy(np.arange(6, dtype=np.float64) + 0.5)
    assert hyd.csv_format(d, 3) == b"0.5,1.5,2.5\n3.5,4.5,5.5\n"


def test_round_trip_full_size(hyd, orc):
    # 2^27 fp64 values (the map chain's array): format, parse, same bits;
    # the text's first and last 100k values against the reference writer
    n = 1 << 27
    print("round trip n=%d" % n, flush=True)
    x = hyd.splitmix_f64(n, 0x435356, 0, symmetric=True)
    cap = hyd.lib().hydhip_csv_text_bound(n)
    text = hyd.DeviceArray((cap,), np.uint8)
    nb = ctypes.c_int64()
    ex = ctypes.c_int64()
    hyd._check(hyd.lib().hydhip_csv_format_f64(x.ptr, n, n, text.ptr, cap, ctypes.byref(nb), ctypes.byref(ex), None))
    y = hyd.DeviceArray((n,), np.float64)
    st = hyd.CsvStatus()
    hyd._check(hyd.lib().hydhip_csv_parse_f64(text.ptr, nb.value, n, y.ptr, ctypes.byref(st), None))
    assert st.error_kind == 0 and st.numbers_found == n
    k = 100_000
    xs = orc.splitmix_f64(k, 0x435356, 0, symmetric=True)
    head = orc.csv_write(xs, 1, k, True)[:-1]
    assert hyd.copy_out(text.ptr, len(head), np.uint8).tobytes() == head
    xe = orc.splitmix_f64(k, 0x435356, n - k, symmetric=True)
    tail = orc.csv_write(xe, 1, k, True)
    assert hyd.copy_out(text.ptr + nb.value - len(tail), len(tail), np.uint8).tobytes() == tail
    # every value back bit for bit: compare on the device via a difference count
    got = y.numpy()
    ref = x.numpy()
    assert (bits(got) == bits(ref)).all()
    for a in (x, y, text):
        a.free()


def test_drop_in_file_write_large(hyd, orc, tmp_path):
    # the text goes out through the pinned halves, one pwrite per half (this
    # text spans two 256 MiB halves, the second partial), also to a device
    # file such as /dev/null
    n = 14_000_003
    print("drop-in file write n=%d" % n, flush=True)
    v = orc.splitmix_f64(n, 88, symmetric=True)
    arr = hyd.host_array(v)
    path = str(tmp_path / "big.csv").encode()
    assert hyd.lib().hyd_write_array_csv_float(arr, path) == 0
    dv = hyd.DeviceArray.from_numpy(v)
    expect = hyd.csv_format(dv)
    with open(path, "rb") as f:
        got = f.read()
    assert len(got) == len(expect) > 256 << 20 and got == expect
    assert hyd.lib().hyd_write_array_csv_float(arr, b"/dev/null") == 0
    hyd.lib().hyd_array_free(arr)
    dv.free()


def test_drop_in_files(hyd, orc, tmp_path):
    rng = np.random.default_rng(8)
    v = random_doubles(rng, 12_000).reshape(30, 400)
    arr = hyd.host_array(v)
    path = str(tmp_path / "out.csv").encode()
    assert hyd.lib().hyd_write_array_csv_float(arr, path) == 0
    with open(path, "rb") as f:
        assert f.read() == orc.csv_write(v.ravel(), 30, 400, True)
    back = hyd.lib().hyd_read_float_array_csv(path, hyd.hyd_tuple(30, 400))
    with open(path, "rb") as f:
        ref = orc.csv_read(f.read(), v.size, True)[0]
    assert (bits(hyd.read_array(back, np.float64)).ravel() == bits(ref)).all()
    i = rng.integers(-2**40, 2**40, (7, 9))
    arr_i = hyd.host_array(i)
    assert hyd.lib().hyd_write_array_csv(arr_i, path) == 0
    with open(path, "rb") as f:
        assert f.read() == orc.csv_write(i.ravel(), 7, 9, False)
    back_i = hyd.lib().hyd_read_array_csv(path, hyd.hyd_tuple(63))
    assert (hyd.read_array(back_i, np.int64) == i.ravel()).all()
    e = hyd.host_array(np.zeros((3, 0)))
    assert hyd.lib().hyd_write_array_csv_float(e, path) == 0
    with open(path, "rb") as f:
        assert f.read() == b"\n\n\n"
    for a in (arr, back, arr_i, back_i, e):
        hyd.lib().hyd_array_free(a)
    for name, count, is_float in (("data.csv", 6, False), ("matA_3x3.csv", 9, True)):
        a = (hyd.lib().hyd_read_float_array_csv if is_float else hyd.lib().hyd_read_array_csv)(
            os.path.join(GOLD, name).encode(), hyd.hyd_tuple(count))
        with open(os.path.join(GOLD, name), "rb") as f:
            ref = orc.csv_read(f.read(), count, is_float)[0]
        assert same(hyd.read_array(a, np.float64 if is_float else np.int64), ref, is_float)
        hyd.lib().hyd_array_free(a)
