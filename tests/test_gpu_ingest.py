"""GPU parity of data ingest: text parsed on the GPU equals the oracle's
restatement of the reference's fscanf loader (input/File.h:233-295) bit for
bit, including lines the device hands to the host; column files load
unchanged; malformed files are rejected."""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle
import topk_amd
from topk_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)
    return str(path)


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _parity(path, n, d, expect_slow=None):
    got, rep = topk_amd.read_csv(path, n, d)
    want, r = oracle.csv_read(path, n, d)
    assert r == n
    assert _same(got, want), np.argwhere(got.view(np.uint32) != want.view(np.uint32))[:5]
    assert rep["rows"] == n and rep["fields"] == d
    if expect_slow is not None:
        assert expect_slow(rep["slow_lines"]), rep
    return got, rep


def test_printf_file_parses_on_device(gpu, tmp_path):
    """2^20 rows of "%f" values (the reference's writer): every line on the
    device path, lines crossing the 64 KiB tiles of the newline scan."""
    n, d = 1 << 20, 8
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 11)
    p = str(tmp_path / "x.csv")
    topk_amd.write_csv(p, x)
    got, rep = _parity(p, n, d, expect_slow=lambda s: s == 0)
    assert rep["lines"] == n and rep["bytes"] == os.path.getsize(p)
    # "%f" keeps 6 decimals: within half a unit of the 6th decimal of the source
    assert np.max(np.abs(got - x)) <= 5e-7 + 1e-7
    rm, _ = topk_amd.read_csv(p, n, d, rowmajor=True)
    assert _same(rm, np.ascontiguousarray(got.T))
    part, _ = topk_amd.read_csv(p, 1000, d)      # the first n rows only
    assert _same(part, got[:, :1000])


@pytest.mark.parametrize("style", ["g9", "e", "g17", "long"])
def test_number_styles(gpu, tmp_path, style):
    rng = np.random.default_rng(7)
    n, d = 3000, 6
    v = rng.uniform(-5, 5, (n, d))
    if style == "g9":
        rows = [",".join("%.9g" % np.float32(a) for a in r) for r in v]
        slow = lambda s: s == 0
    elif style == "e":
        rows = [",".join("%e" % (a * 10.0 ** rng.integers(-12, 12)) for a in r) for r in v]
        slow = lambda s: s == 0
    elif style == "g17":
        rows = [",".join("%.17g" % a for a in r) for r in v]
        slow = lambda s: s >= 0.9 * n          # 17 significant digits: the host converts them
    else:
        rows = [",".join("%.30f" % a for a in r) for r in v]
        slow = lambda s: s == n
    p = _write(tmp_path / "s.csv", "\n".join(rows) + "\n")
    got, _ = _parity(p, n, d, expect_slow=slow)
    if style == "g9":
        assert _same(got, v.T.astype(np.float32))  # 9 digits round-trip binary32


def _midpoint_decimals(rng, count):
    """15-digit decimals whose binary64 rounding is exactly a binary32
    midpoint although they are not one (double rounding would go wrong)."""
    out = []
    while len(out) < count:
        f = np.float32(rng.uniform(1, 2)) * np.float32(2.0 ** int(rng.integers(-20, 20)))
        mid = Fraction(float(f)) + Fraction(float(np.spacing(f))) / 2
        s = "%.15g" % float(mid)
        if Fraction(s) != mid and float(s) == float(mid):
            out.append(s)
    return out


def test_binary32_midpoints_go_to_host(gpu, tmp_path):
    rng = np.random.default_rng(5)
    vals = _midpoint_decimals(rng, 64)
    lines = [",".join(vals[i:i + 4]) for i in range(0, 64, 4)]
    filler = ["0.25,0.5,0.75,1.0"] * 16
    text = "\n".join(a for pair in zip(filler, lines) for a in pair) + "\n"
    p = _write(tmp_path / "m.csv", text)
    _parity(p, 32, 4, expect_slow=lambda s: s == 16)


def test_layout_variants(gpu, tmp_path):
    # blank lines first; sign, negative zero, bare fraction / integer, CRLF; leading tab,
    # trailing blanks; leading zeros, subnormal, large exponent, overflow, hex, no final newline
    text = "\n  \n+1.5,-0.000000,.5,5.\r\n 2e3,\t7E-2,3,1e+1  \n\n0,00012,1e-40,3e38\n-4.25,1e39,0x1p-3,9"
    p = _write(tmp_path / "v.csv", text)
    assert topk_amd.csv_count(p) == (4, 4)
    got, rep = _parity(p, 4, 4, expect_slow=lambda s: s == 2)   # subnormal / overflow / hex lines
    assert rep["lines"] == 7
    assert got[1, 0].tobytes() == np.float32(-0.0).tobytes()
    assert np.isinf(got[1, 3])


def test_single_values_and_long_lines(gpu, tmp_path):
    p = _write(tmp_path / "one.csv", "0.125")             # one line, no newline, d = 1
    got, _ = _parity(p, 1, 1)
    assert got[0, 0] == 0.125
    rng = np.random.default_rng(9)
    long_rows = [",".join("%.40f" % a for a in rng.uniform(0, 1, 32)) for _ in range(300)]
    p2 = _write(tmp_path / "long.csv", "\n".join(long_rows) + "\n")
    _parity(p2, 300, 32)


def test_nonfinite_text(gpu, tmp_path):
    p = _write(tmp_path / "nan.csv", "1,nan\ninf,-inf\n")
    got, rep = _parity(p, 2, 2)
    assert np.isnan(got[1, 0]) and np.isinf(got[0, 1])
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.Engine.from_file(p)
    assert e.value.status == _lib.TOPK_ERR_NONFINITE


@pytest.mark.parametrize("text", ["1,2,3\n4,5\n", "1,2,3\n4,5,6,7\n", "1,2,3\n4 ,5,6\n", "1,,3\n",
                                  "1,2,3\n4,x,6\n", "1,2,3\n4,5,6,\n", "1,2,3\n1e,2,3\n"])
def test_malformed_rows_are_errors(gpu, tmp_path, text):
    p = _write(tmp_path / "bad.csv", text)
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.read_csv(p, 0, 3)
    assert e.value.status == _lib.TOPK_ERR_FORMAT


def test_shape_errors(gpu, tmp_path):
    p = _write(tmp_path / "a.csv", "1,2\n3,4\n")
    for n, d in ((3, 2), (2, 3)):
        with pytest.raises(topk_amd.TopkError) as e:
            topk_amd.read_csv(p, n, d)
        assert e.value.status == _lib.TOPK_ERR_FORMAT
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.Engine.from_file(str(tmp_path / "missing.csv"))
    assert e.value.status == _lib.TOPK_ERR_IO


def test_engine_from_text_answers_like_the_table(gpu, tmp_path):
    n, d = 200003, 8
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 3)
    p = str(tmp_path / "a.csv")
    topk_amd.write_csv(p, x)
    want, _ = oracle.csv_read(p, n, d)
    with topk_amd.Engine.from_file(p) as e:
        assert e.n == n and e.d == d and e.ingest["format"] == 0 and e.ingest["slow_lines"] == 0
        assert _same(e.read_columns(), want)
        w = np.linspace(0.2, 1.0, d).astype(np.float32)
        ids, sc, thr = e.fte(64, w, list(range(d)))
    ri, rs, rt = oracle.fte_topk(want, n, 64, w, list(range(d)))
    assert np.array_equal(ids, ri) and _same(sc, rs) and thr == rt


def test_column_file_loads_unchanged(gpu, tmp_path):
    n, d = 300001, 6
    x = topk_amd.generate(n, d, topk_amd.MIXED, 4)
    p = str(tmp_path / "x.col")
    topk_amd.write_colfile(p, x)
    with topk_amd.Engine.from_file(p) as e:
        assert e.ingest["format"] == 1 and e.ingest["rows"] == n
        assert _same(e.read_columns(), x)
    with topk_amd.Engine.from_file(p, n=1000, d=4) as e:      # a prefix of the objects and attributes
        assert _same(e.read_columns(), np.ascontiguousarray(x[:4, :1000]))
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.Engine.from_file(p, n=n + 1)
    assert e.value.status == _lib.TOPK_ERR_FORMAT


def test_cpp_driver_loads_text(gpu, tmp_path):
    n, d = 100000, 6
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0)
    p = str(tmp_path / "d.csv")
    topk_amd.write_csv(p, x)
    exe = os.path.join(ROOT, "topk_amd", "bin", "bench_topk")
    r = subprocess.run([exe, "-n", str(n), "-d", str(d), "-f", p, "-s", "16", "-e", "32", "--check"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Loading data from file" in r.stdout and "FAILED" not in r.stdout
    assert r.stdout.count("PASSED") == 2 * 5
