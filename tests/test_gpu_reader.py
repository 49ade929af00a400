"""Reader.get_chunk over a row-group directory: every vector of every column equals the written values, through the
Python mirror and through the C++ Reader (include/fls_reader.hpp) in a test program built by the library's
Makefile (tests/cpp/read_fls_dir.cpp)."""
import os
import subprocess

import numpy as np
import pytest

from fastlanes_amd import workloads as wl
from fastlanes_amd.reader import Connection

pytestmark = pytest.mark.gpu


def test_reader_get_chunk(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    n = 40 * 1024
    cols = [wl.rowgroup_column("bitpack", 8, n, 1), wl.rowgroup_column("delta", 64, n, 2),
            wl.rowgroup_column("rle", 32, n, 3), wl.rowgroup_column("dict", 16, n, 4),
            wl.geometric_runs(n, 6.0, seed=5, value_bits=10).astype(np.int64), np.linspace(0, 1, n),
            wl.dict_runs(n, np.uint16, seed=6)]
    exprs = [1, 11, 6, 17, 7, 0, 27]
    Connection.to_fls(str(tmp_path), cols, exprs)
    rd = Connection().read_fls(str(tmp_path))
    assert rd.n_vec == 40
    for v in (0, 17, 39):
        ch = rd.get_chunk(v)
        assert len(ch) == len(cols)
        for c, x in enumerate(cols):
            np.testing.assert_array_equal(ch[c].cpu().numpy(), x[v * 1024:(v + 1) * 1024], err_msg=f"col {c} vec {v}")
    for c, x in enumerate(cols):
        np.testing.assert_array_equal(rd.column(c), x)
    with pytest.raises(IndexError):
        rd.get_chunk(40)


def _columns(n):
    return [wl.rowgroup_column("bitpack", 8, n, 1), wl.rowgroup_column("delta", 64, n, 2),
            wl.rowgroup_column("rle", 32, n, 3), wl.rowgroup_column("dict", 16, n, 4),
            wl.geometric_runs(n, 6.0, seed=5, value_bits=10).astype(np.int64), np.linspace(0, 1, n),
            wl.dict_runs(n, np.uint16, seed=6), wl.zipf_codes(n, 4096, 1.1, seed=7, dtype=np.int32),
            (np.arange(n) % 77).astype(np.int8)]


def test_cpp_reader_program(tmp_path):
    """fastlanes::Connection::read_fls + Reader::get_chunk in C++: every vector of every column bit-exact (twice,
    around reset()), get_chunk past the end throws, and to_csv writes every row."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = os.path.join(os.path.dirname(os.path.abspath(wl.__file__)), "bin", "read_fls_dir")
    assert os.path.exists(exe), "build the library first (make -C fastlanes_amd)"
    n = 33 * 1024
    cols = _columns(n)
    exprs = [1, 11, 6, 17, 7, 0, 27, 22, 6]
    Connection.to_fls(str(tmp_path), cols, exprs)
    with open(os.path.join(tmp_path, "expected.bin"), "wb") as f:
        for x in cols:
            f.write(np.ascontiguousarray(x).tobytes())
    print("C++ reader program", flush=True)
    r = subprocess.run([exe, str(tmp_path), os.path.join(tmp_path, "expected.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == f"ok 33 {len(cols)}", r.stdout + r.stderr
    rows = open(os.path.join(tmp_path, "extracted_from_fls.csv")).read().splitlines()
    assert len(rows) == n
    for i in (0, 1, 1023, 1024, n - 1):
        fields = rows[i].split("|")
        assert len(fields) == len(cols)
        for c, x in enumerate(cols):
            want = float(x[i]) if x.dtype.kind == "f" else int(x[i])
            got = float(fields[c]) if x.dtype.kind == "f" else int(fields[c])
            assert got == want, (i, c, fields[c], x[i])

