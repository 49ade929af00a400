"""DICT_RLE pages (pool id 27: DICT_RLE / DELTA_NO_PATCH_REC, src/cor/exp_pool.cpp:53) in the oracle and the host
encoder. No reference test or fixture holds such a page; the layout is checked against the pinned RLE and
dictionary pages it is built from, by hand on a small page, and by round trips."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

DTYPES = [np.uint8, np.int16, np.uint32, np.int32, np.int64, np.uint64]


def _page(enc: np.ndarray, pos: int, arr_c: int):
    """(header, offsets, entry points, buffers) of the data page at pos (last page of the column)."""
    hdr = enc[pos:pos + 8]
    vec_c = int(hdr[4]) | int(hdr[5]) << 8
    off = np.frombuffer(enc[pos + 8:pos + 8 + 4 * arr_c].tobytes(), np.uint32)
    eps = enc[pos + 8 + 4 * arr_c:pos + off[0]]
    bufs = [enc[pos + off[b]:pos + (off[b + 1] if b + 1 < arr_c else enc.size - pos)] for b in range(arr_c)]
    return hdr, off, eps, bufs, vec_c


def _dict_page_size(enc: np.ndarray) -> int:
    """Bytes of the dictionary page at the start of enc (64-bit values: one HZL tail chunk for < 1024 entries)."""
    n = int(np.frombuffer(enc[4:8].tobytes(), np.uint32)[0])
    assert n < 1024
    assert enc[8] == 255
    return 8 + 12 + int(np.frombuffer(enc[12:20].tobytes(), np.uint64)[0])


def test_small_page_by_hand(oracle):
    """Three runs in one vector: dictionary {7, 5} in first-occurrence order, run codes [0, 1, 0] framed as one HZL
    tail chunk, positions as in an RLE page of the codes."""
    x = np.concatenate([np.full(100, 7), np.full(900, 5), np.full(24, 7)]).astype(np.int64)
    enc = oracle.encode_column(x, 27)
    # dictionary page: {exp_t, 0, 0, 0, entry_c} + tail chunk {bw 255, 0, exc_c 0, byte count} + 7, 5
    assert enc[0] == 8 and int(np.frombuffer(enc[4:8].tobytes(), np.uint32)[0]) == 2
    assert list(enc[8:12]) == [255, 0, 0, 0] and int(np.frombuffer(enc[12:20].tobytes(), np.uint64)[0]) == 16
    assert list(np.frombuffer(enc[20:36].tobytes(), np.int64)) == [7, 5]
    hdr, off, eps, bufs, vec_c = _page(enc, 36, 5)
    assert list(hdr) == [11, 4, 0, 0, 1, 0, 0, 0] and vec_c == 1 and eps.size == 0
    assert list(off) == [28, 28 + 24, 28 + 24 + 128, 28 + 24 + 128 + 1, 28 + 24 + 128 + 1 + 2]
    run = bufs[0]
    assert list(run[:4]) == [255, 0, 0, 0] and int(np.frombuffer(run[4:12].tobytes(), np.uint64)[0]) == 12
    assert list(np.frombuffer(run[12:24].tobytes(), np.uint32)) == [0, 1, 0]
    # positions: those of an RLE DELTA_NO_PATCH page (pool id 6) of the codes column
    codes = np.concatenate([np.zeros(100), np.ones(900), np.zeros(24)]).astype(np.uint32)
    rle = oracle.encode_column(codes, 6)
    _, _, _, rbufs, _ = _page(rle, 0, 5)
    assert list(np.frombuffer(rbufs[0].tobytes(), np.uint32)) == [0, 1, 0]
    for b in range(1, 5):
        np.testing.assert_array_equal(bufs[b], rbufs[b])
    # positions 0..2 rise by at most one per step: bw 1 (delta), base 0
    assert bufs[2][0] == 1 and list(bufs[3]) == [0, 0]
    np.testing.assert_array_equal(oracle.decode_column(enc, np.int64, x.size), x)


def test_page_is_rle_page_of_codes_with_hzl_run_buffer(oracle):
    """Many vectors (run codes span several 1024-code HZL chunks): entry points and position buffers equal the RLE
    page of the code column; buffer 0 decodes (HZL) to the RLE page's run values; the dictionary page equals the
    DICT page's."""
    x = wl.dict_runs(40 * 1024, np.int64, n_keys=500, seed=3)
    enc = oracle.encode_column(x, 27)
    # dictionary codes in first-occurrence order (DICT, dict_fac.cpp:259-330)
    uniq, first = np.unique(x, return_index=True)
    rank = np.empty(uniq.size, np.uint32)
    rank[np.argsort(first)] = np.arange(uniq.size, dtype=np.uint32)
    codes = rank[np.searchsorted(uniq, x)]
    dsz = _dict_page_size(enc)
    dict_enc = oracle.encode_column(x, 16)
    np.testing.assert_array_equal(enc[:dsz], dict_enc[:dsz])
    rle = oracle.encode_column(codes, 6)
    hdr, off, eps, bufs, vec_c = _page(enc, dsz, 5)
    rhdr, roff, reps, rbufs, rvec_c = _page(rle, 0, 5)
    assert vec_c == rvec_c == 40 and hdr[0] == 11 and hdr[1] == 4
    np.testing.assert_array_equal(eps, reps)
    for b in range(1, 5):
        np.testing.assert_array_equal(bufs[b], rbufs[b])
    runs = np.frombuffer(rbufs[0].tobytes(), np.uint32)
    assert runs.size > 2048  # several FFOR chunks and a tail
    dec, used = oracle.hzl_decode(np.ascontiguousarray(bufs[0]), runs.size, np.uint32)
    assert dec.size == runs.size and used == bufs[0].size
    np.testing.assert_array_equal(dec, runs)
    np.testing.assert_array_equal(oracle.decode_column(enc, np.int64, x.size), x)


@pytest.mark.parametrize("dtype", DTYPES)
def test_round_trip_and_encoders_agree(oracle, dtype):
    x = wl.dict_runs(24 * 1024, dtype, seed=np.dtype(dtype).itemsize)
    enc = oracle.encode_column(x, 27)
    np.testing.assert_array_equal(oracle.decode_column(enc, dtype, x.size), x)
    np.testing.assert_array_equal(fa.encode_from_memory(x, 27), enc)


def test_every_value_its_own_run(oracle):
    """1024 runs per vector (the position bit width is 1 everywhere) and a dictionary larger than one chunk."""
    x = np.arange(8 * 1024, dtype=np.int64) * 3 + 11
    enc = oracle.encode_column(x, 27)
    np.testing.assert_array_equal(oracle.decode_column(enc, np.int64, x.size), x)
    np.testing.assert_array_equal(fa.encode_from_memory(x, 27), enc)


def test_empty_column(oracle):
    """No vectors: an empty dictionary page and a data page whose buffers are all empty."""
    x = np.zeros(0, np.int64)
    enc = oracle.encode_column(x, 27)
    assert enc.size == 20 + 8 + 4 * 5  # dictionary page (tail chunk of 0 bytes), page header and offsets
    np.testing.assert_array_equal(fa.encode_from_memory(x, 27), enc)
    assert oracle.decode_column(enc, np.int64, 0).size == 0


def test_rejected_for_doubles_and_bad_codes(oracle):
    with pytest.raises(ValueError):
        oracle.encode_column(np.zeros(1024, np.float64), 27)
    x = np.concatenate([np.full(512, 3), np.full(512, 4)]).astype(np.int64)
    enc = oracle.encode_column(x, 27)
    bad = enc.copy()
    dsz = _dict_page_size(enc)
    _, off, _, _, _ = _page(enc, dsz, 5)
    bad[dsz + off[0] + 12: dsz + off[0] + 16] = np.frombuffer(np.uint32(9).tobytes(), np.uint8)  # code 9 of 2
    with pytest.raises(ValueError):
        oracle.decode_column(bad, np.int64, x.size)


def test_random_access_reports_no_buffers():
    """No vector owns a slice of the page-wide run-code chain: init_random_access reports zero buffers."""
    import ctypes as C

    x = wl.dict_runs(4 * 1024, np.int64, seed=12)
    enc = fa.encode_from_memory(x, 27)
    nb = C.c_uint64(99)
    pairs = (C.c_uint64 * 16)()
    fa.lib().init_random_access(enc.ctypes.data, 1, C.byref(nb), pairs)
    assert nb.value == 0
