"""Data ingest on the CPU: the oracle's restatement of the reference's text
loader (File<T>::count / read_scanf_t, input/File.h:107-128,233-295) pinned
against exact rational arithmetic, the library's host-side ingest calls
(row count, text and column-file writers, column-file header), and the
argument checks of the host-memory store."""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
import topk_amd
from topk_amd import _lib


def nearest_f32(text):
    """Correctly rounded (ties to even) binary32 of a decimal string, from
    exact rational arithmetic."""
    t = text.strip()
    fr = Fraction(t)
    neg = t.startswith("-")
    if fr == 0:
        return np.float32(-0.0) if neg else np.float32(0.0)
    a = abs(fr)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)
    ulp = Fraction(2) ** (e - 23)
    q = a / ulp
    m = q.numerator // q.denominator
    r = q - m
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and m % 2 == 1):
        m += 1
    v = m * ulp
    if v >= Fraction(2) ** 128:
        f = np.float32(np.inf)
    else:
        f = np.float32(float(v))  # m <= 2^24 times a power of two: exact in binary64 and binary32
    return -f if fr < 0 else f


def test_nearest_f32_known_answers():
    assert nearest_f32("0.5") == np.float32(0.5)
    assert nearest_f32("-0.0").tobytes() == np.float32(-0.0).tobytes()
    # 1 + 2^-24 is the midpoint between 1 and 1 + 2^-23: ties to even gives 1
    assert nearest_f32("1.000000059604644775390625") == np.float32(1.0)
    # 1 + 3*2^-24 is the midpoint between 1 + 2^-23 and 1 + 2^-22: ties to even gives 1 + 2^-22
    assert nearest_f32("1.000000178813934326171875") == np.float32(1 + 2 ** -22)
    assert nearest_f32("1.00000017881393433") == np.float32(1 + 2 ** -22)  # just above: up
    assert nearest_f32("1.0000001788139343") == np.float32(1 + 2 ** -23)   # just below: down
    assert nearest_f32("1.0000000596046448") == np.float32(1 + 2 ** -23)   # just above the first midpoint
    assert nearest_f32("1.0000000596046447") == np.float32(1.0)            # just below it
    assert nearest_f32("3.4028235677973366e38") == np.float32(3.4028234663852886e38)
    assert np.isinf(nearest_f32("3.5e38"))
    assert nearest_f32("1e-45") == np.float32(1.401298464324817e-45)      # smallest subnormal
    rng = np.random.default_rng(0)
    for v in rng.standard_normal(200).astype(np.float32):
        assert nearest_f32(repr(float(v))) == v                            # binary32 values are exact decimals


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)
    return path


def _random_strings(rng, rows, d, style):
    vals = rng.uniform(-3, 3, (rows, d))
    out = []
    for row in vals:
        if style == "f":
            out.append(",".join("%f" % v for v in row))
        elif style == "g9":
            out.append(",".join("%.9g" % np.float32(v) for v in row))
        elif style == "e":
            out.append(",".join("%e" % (v * 10.0 ** rng.integers(-30, 30)) for v in row))
        elif style == "long":
            out.append(",".join("%.25f" % v for v in row))
        else:
            raise ValueError(style)
    return out


@pytest.mark.parametrize("style", ["f", "g9", "e", "long"])
def test_oracle_csv_read_is_correctly_rounded(tmp_path, style):
    rng = np.random.default_rng(len(style))
    rows, d = 300, 5
    lines = _random_strings(rng, rows, d, style)
    p = _write(tmp_path / "t.csv", "\n".join(lines) + "\n")
    got, r = oracle.csv_read(str(p), rows, d)
    assert r == rows
    want = np.array([[nearest_f32(s) for s in ln.split(",")] for ln in lines], np.float32).T
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # row-major layout holds the same values
    rm, _ = oracle.csv_read(str(p), rows, d, colmajor=False)
    assert np.array_equal(rm, got.T)


def test_oracle_csv_count_and_stop_rules(tmp_path):
    p = _write(tmp_path / "a.csv", "1,2,3\n4,5,6\n7,8,9\n")
    assert oracle.csv_count(str(p)) == (3, 3)
    out, r = oracle.csv_read(str(p), 2, 3)          # the first n rows only
    assert r == 2 and out[:, 1].tolist() == [4, 5, 6]
    q = _write(tmp_path / "b.csv", "1,2\n\n  3,4\r\n5,6")   # blank line skipped, CRLF, no final newline
    assert oracle.csv_count(str(q)) == (3, 2)                # newlines, as File::count counts them
    out, r = oracle.csv_read(str(q), 3, 2)
    assert r == 3 and out.tolist() == [[1, 3, 5], [2, 4, 6]]
    s = _write(tmp_path / "c.csv", "1,2,3\n4,5\n")
    with pytest.raises(ValueError):
        oracle.csv_read(str(s), 2, 3)


def test_library_csv_count(tmp_path):
    p = _write(tmp_path / "a.csv", "1,2,3\n4,5,6\n7,8,9\n")
    assert topk_amd.csv_count(str(p)) == oracle.csv_count(str(p)) == (3, 3)
    # blank lines are not rows; a last line without a newline is one
    q = _write(tmp_path / "b.csv", "\n \n1,2\n\n3,4\r\n5,6")
    assert topk_amd.csv_count(str(q)) == (3, 2)
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.csv_count(str(tmp_path / "missing.csv"))
    assert e.value.status == _lib.TOPK_ERR_IO


def test_csv_writer_matches_printf(tmp_path):
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 2, (3, 1000)).astype(np.float32)
    x[0, :4] = [0.0, -0.0, 1e-7, 123456.75]
    p = str(tmp_path / "w.csv")
    topk_amd.write_csv(p, x)
    want = "".join(",".join("%f" % float(v) for v in x[:, i]) + "\n" for i in range(x.shape[1]))
    assert open(p).read() == want
    # row-major input writes the same text
    p2 = str(tmp_path / "w2.csv")
    topk_amd.write_csv(p2, np.ascontiguousarray(x.T), rowmajor=True)
    assert open(p2).read() == want
    # and the oracle reads back the correctly rounded "%f" values
    got, r = oracle.csv_read(p, 1000, 3)
    assert r == 1000
    assert np.array_equal(got, np.array([[nearest_f32("%f" % float(v)) for v in col] for col in x], np.float32))


def test_colfile_roundtrip_on_host(tmp_path):
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (5, 777)).astype(np.float32)
    p = str(tmp_path / "x.col")
    topk_amd.write_colfile(p, x)
    assert topk_amd.colfile_info(p) == (777, 5)
    raw = open(p, "rb").read()
    assert raw[:8] == b"TOPKCOL1" and len(raw) == 4096 + x.nbytes
    assert np.array_equal(np.frombuffer(raw, np.float32, offset=4096).reshape(5, 777), x)
    p2 = str(tmp_path / "y.col")
    topk_amd.write_colfile(p2, np.ascontiguousarray(x.T), rowmajor=True)
    assert open(p2, "rb").read() == raw
    bad = _write(tmp_path / "z.col", "1,2\n")
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.colfile_info(str(bad))
    assert e.value.status == _lib.TOPK_ERR_FORMAT


def test_text_read_needs_a_gpu(tmp_path):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    p = _write(tmp_path / "a.csv", "1,2\n")
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.read_csv(str(p))
    assert e.value.status == _lib.TOPK_ERR_NO_DEVICE
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.Engine.from_file(str(p))
    assert e.value.status == _lib.TOPK_ERR_NO_DEVICE


def test_host_store_errors_without_a_gpu():
    from conftest import gpu_available
    L = topk_amd.lib()
    import ctypes
    h = ctypes.c_void_p()
    assert L.topk_hstore_create(ctypes.byref(h), 0, 4, 0) == _lib.TOPK_ERR_INVALID      # n = 0
    assert L.topk_hstore_create(ctypes.byref(h), 100, 33, 0) == _lib.TOPK_ERR_INVALID    # d > 32
    assert L.topk_hstore_destroy(None) == _lib.TOPK_ERR_INVALID
    assert L.topk_hstore_fte(None, 4, 1, None, None, None, None, None) == _lib.TOPK_ERR_INVALID
    if not gpu_available():
        with pytest.raises(topk_amd.TopkError) as e:
            topk_amd.HostStore(1000, 4)
        assert e.value.status == _lib.TOPK_ERR_NO_DEVICE
