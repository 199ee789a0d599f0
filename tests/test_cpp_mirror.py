"""The C++ mirror of the reference's objects (include/topk_amd.hpp) on the
CPU: a caller written like validation/bench_cpu.h builds against it, and
its File object writes and counts text files with the reference's format
(input/File.h:107-128 count, :415-430 write_printf_t) without a GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from topk_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include "topk_amd.hpp"

int main(int argc, char **argv) {
    const uint64_t n = 5, d = 3;
    float *data = nullptr;
    topk_amd::File g(argv[1], false, n, d);
    g.set_transpose(true);
    g.gen(data, 1);                       // column-major synthetic table
    g.store(argv[1], data);               // "%f" rows
    topk_amd::File f(argv[1], false);     // counts rows and fields
    std::printf("%llu %llu\n", (unsigned long long)f.rows(), (unsigned long long)f.items());
    for (uint64_t i = 0; i < n * d; i++) std::printf("%.9g\n", data[i]);
    std::free(data);
    try {
        topk_amd::File missing(std::string(argv[1]) + ".missing", false);
    } catch (const topk_amd::error &e) {
        std::printf("status %d\n", e.status);
    }
    return 0;
}
"""


def test_file_mirror_writes_and_counts(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    src = tmp_path / "caller.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "caller"
    lib_dir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", lib_dir, "-ltopk_amd", "-Wl,-rpath," + lib_dir], check=True, capture_output=True)
    out_file = tmp_path / "d_5_3_i"
    r = subprocess.run([str(exe), str(out_file)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[0] == "5 3"
    vals = np.array([float(v) for v in lines[1:16]], np.float32).reshape(3, 5)   # column-major x[m*n+i]
    import topk_amd
    assert np.array_equal(vals, topk_amd.generate(5, 3, topk_amd.INDEPENDENT, 0))
    want = "".join(",".join("%f" % float(v) for v in vals[:, i]) + "\n" for i in range(5))
    assert out_file.read_text() == want
    assert lines[16] == "status %d" % _lib.TOPK_ERR_IO
