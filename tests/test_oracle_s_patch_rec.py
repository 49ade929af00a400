"""S_PATCH_REC pages (pool id 3, src/cor/exp_pool.cpp:26; RPN include/fls/cor/exp/ffor/ffor_exp.hpp:87-94): S_PATCH
vectors whose page-wide exception values are one HZL chain of u64 values. No reference test or fixture holds such a
page; the layout is checked against the pinned S_PATCH page it is built from, and by round trips."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl


def _page(enc: np.ndarray, arr_c: int = 5):
    hdr = enc[:8]
    off = np.frombuffer(enc[8:8 + 4 * arr_c].tobytes(), np.uint32)
    eps = enc[8 + 4 * arr_c:off[0]]
    bufs = [enc[off[b]:(off[b + 1] if b + 1 < arr_c else enc.size)] for b in range(arr_c)]
    return hdr, eps, bufs


@pytest.mark.parametrize("dtype", [np.int64, np.uint64])
def test_page_is_s_patch_page_with_hzl_exceptions(oracle, dtype):
    x = wl.patch_column(dtype, 200 * 1024, seed=3)  # ~2000 exceptions: one full HZL chunk and a tail
    rec, s = oracle.encode_column(x, 3), oracle.encode_column(x, 2)
    hdr, eps, bufs = _page(rec)
    shdr, seps, sbufs = _page(s)
    assert hdr[0] == 4 and hdr[1] == 5 and shdr[1] == 3 and list(hdr[2:]) == list(shdr[2:])
    np.testing.assert_array_equal(eps, seps)
    for b in range(1, 5):
        np.testing.assert_array_equal(bufs[b], sbufs[b])
    exc = np.frombuffer(sbufs[0].tobytes(), np.uint64)
    assert exc.size > 1024
    dec, used = oracle.hzl_decode(np.ascontiguousarray(bufs[0]), exc.size, np.uint64)
    assert used == bufs[0].size
    np.testing.assert_array_equal(dec, exc)
    np.testing.assert_array_equal(oracle.decode_column(rec, dtype, x.size), x)
    np.testing.assert_array_equal(fa.encode_from_memory(x, 3), rec)


def test_no_exceptions_and_rejections(oracle):
    x = (np.arange(4 * 1024) % 100).astype(np.int64)  # no outliers: an empty chain (one tail header)
    rec = oracle.encode_column(x, 3)
    _, _, bufs = _page(rec)
    assert list(bufs[0]) == [255] + [0] * 11
    np.testing.assert_array_equal(oracle.decode_column(rec, np.int64, x.size), x)
    np.testing.assert_array_equal(fa.encode_from_memory(x, 3), rec)
    for dt in (np.uint32, np.int32, np.float64):  # the reference's pool holds it for int64 only; 64-bit here
        with pytest.raises(ValueError):
            oracle.encode_column(np.zeros(1024, dt), 3)
