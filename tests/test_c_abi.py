"""CPU-side checks of the C-ABI library: it loads, exports every declared symbol, and its host encoder writes
the same bytes as the oracle's encoder (no GPU needed for any of these)."""
import os
import re

import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("fastlanes.h", "fls_gpu.h", "fls_wizard.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_library_exports_every_declared_symbol():
    lib = fa.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


DT = [np.int64, np.uint32, np.int32, np.uint16, np.uint8, np.uint64, np.int16, np.int8]


@pytest.mark.parametrize("dtype", DT)
@pytest.mark.parametrize("eid", [0, 1, 6, 7, 11, 16, 17, 22, 28, 29])
def test_host_encoder_matches_oracle_bytes(oracle, dtype, eid):
    g = np.random.default_rng(100 + eid)
    info = np.iinfo(dtype)
    n = 4 * 1024
    if eid in (6, 7):
        x = wl.geometric_runs(n, 6.0, seed=eid, value_bits=min(20, info.bits - 1)).astype(dtype)
    elif eid >= 16:
        keys = g.integers(info.min, info.max, 200, dtype=dtype, endpoint=True)
        x = keys[g.integers(0, 200, n)]
    else:
        x = g.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    mine = fa.encode_from_memory(x, eid)
    ref = oracle.encode_column(x, eid)
    np.testing.assert_array_equal(mine, ref)


def test_host_encoder_multi_page_matches_oracle(oracle):
    x = wl.timestamps(66_000 * 1024)
    np.testing.assert_array_equal(fa.encode_from_memory(x, 11), oracle.encode_column(x, 11))


def test_encode_errors():
    with pytest.raises(fa.FlsError):
        fa.encode_from_memory(np.arange(1000, dtype=np.int64), 1)  # not a multiple of 1024
    with pytest.raises(fa.FlsError):
        fa.encode_from_memory(np.arange(1024, dtype=np.uint32), 3)  # S_PATCH_REC: 64-bit only
    with pytest.raises(fa.FlsError):
        fa.encode_from_memory(np.arange(1024, dtype=np.uint16), 2)  # patched FFOR: 32/64-bit only
    with pytest.raises(fa.FlsError):
        fa.encode_from_memory(np.arange(2048, dtype=np.int64), 0, capacity=100)  # capacity


def test_plan_rejects_corrupt_pages(oracle):
    import ctypes as C
    x = np.arange(4096, dtype=np.int64)
    enc = oracle.encode_column(x, 1)
    lib = fa.lib()
    h = C.c_void_p()
    bad = enc.copy()
    bad[0] = 99  # unknown expression
    assert lib.fls_gpu_plan_create(bad.ctypes.data, bad.size, 5, C.byref(h)) == -2
    bad = enc.copy()
    bad[8:12] = 255  # buffer offset beyond the page
    assert lib.fls_gpu_plan_create(bad.ctypes.data, bad.size, 5, C.byref(h)) == -2
    assert lib.fls_gpu_plan_create(enc.ctypes.data, enc.size - 10, 5, C.byref(h)) == -2  # truncated
    # a bit width larger than the type
    pos = 8 + 12 + 12 * 3  # first bw byte (vector 0)
    bad = enc.copy()
    bad[pos] = 65
    assert lib.fls_gpu_plan_create(bad.ctypes.data, bad.size, 5, C.byref(h)) == -2
