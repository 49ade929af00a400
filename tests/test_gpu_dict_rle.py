"""DICT_RLE pages (pool id 27) on the GPU: decode and predicates through the C ABI against the CPU oracle, bit-exact."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from tests.test_gpu_decode import _bitmap_words, dev, empty, to_np, torch_cuda  # noqa: F401

pytestmark = pytest.mark.gpu

DTYPES = [np.uint8, np.int16, np.uint32, np.int32, np.int64, np.uint64]


@pytest.mark.parametrize("mean", [1.0, 12.0])
@pytest.mark.parametrize("dtype", DTYPES)
def test_decode_matches_oracle(torch_cuda, oracle, dtype, mean):
    torch = torch_cuda
    print(f"dict_rle dtype={np.dtype(dtype).name} mean={mean}", flush=True)
    x = wl.dict_runs(40 * 1024, dtype, n_keys=3000, mean=mean, seed=np.dtype(dtype).itemsize)
    enc = fa.encode_from_memory(x, 27)
    want = oracle.decode_column(enc, dtype, x.size)
    np.testing.assert_array_equal(want, x)
    plan = fa.Plan(enc, dtype)
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, dtype, x.size), want)


def test_multi_page(torch_cuda, oracle):
    torch = torch_cuda
    x = wl.dict_runs(65_600 * 1024, np.int64, n_keys=4096, seed=9)
    enc = fa.encode_from_memory(x, 27)
    plan = fa.Plan(enc, np.int64)
    assert plan.n_pages == 2
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, np.int64, x.size), x)


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.int64])
def test_predicates_match_numpy(torch_cuda, oracle, dtype):
    torch = torch_cuda
    x = wl.dict_runs(24 * 1024, dtype, seed=5)
    enc = fa.encode_from_memory(x, 27)
    want_x = oracle.decode_column(enc, dtype, x.size)
    plan = fa.Plan(enc, dtype)
    d_in = dev(torch, enc)
    n_vec = x.size // 1024
    for op, fn, c in [("<", np.less, x[777]), ("==", np.equal, x[5000]), (">=", np.greater_equal, x[12345])]:
        d_bm = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
        d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
        plan.select(d_in.data_ptr(), op, dtype(c), d_bm.data_ptr(), d_cnt.data_ptr())
        plan.check()
        mask = fn(want_x, dtype(c))
        np.testing.assert_array_equal(d_bm.cpu().numpy().view(np.uint64), _bitmap_words(mask), err_msg=op)
        np.testing.assert_array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1), err_msg=op)


def test_out_of_range_run_code_is_reported(torch_cuda, oracle):
    torch = torch_cuda
    x = np.concatenate([np.full(512, 3), np.full(512, 4)] * 4).astype(np.int64)
    enc = fa.encode_from_memory(x, 27)
    dsz = 8 + 12 + 16  # dictionary page: two 64-bit entries in one HZL tail chunk
    off0 = int(np.frombuffer(enc[dsz + 8:dsz + 12].tobytes(), np.uint32)[0])
    bad = enc.copy()
    bad[dsz + off0 + 12:dsz + off0 + 16] = np.frombuffer(np.uint32(9).tobytes(), np.uint8)  # first run code: 9 of 2
    plan = fa.Plan(bad, np.int64)
    d_in, d_out = dev(torch, bad), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    with pytest.raises(fa.FlsError):
        plan.check()


@pytest.mark.parametrize("dtype", [np.int64, np.uint32])
def test_dictionary_with_hzl_exceptions(torch_cuda, oracle, dtype):
    """A dictionary chain with exceptions (as the reference's HZL compressor writes them, src/cor/eng/hzl_cmp.cpp;
    layout src/cor/eng/hzl_de_cmp.cpp:40-58) spliced before a DICT_RLE page: the dictionary is decoded by the
    HZL kernels first, then the run values looked up."""
    torch = torch_cuda
    DW = 64 if np.dtype(dtype).itemsize == 8 else 32
    ddt = np.uint64 if DW == 64 else np.uint32
    K = 1000
    v = (np.arange(1024, dtype=np.uint64) * K + 5).astype(dtype)
    x = np.tile(v, 8)  # every value its own run; codes 0..1023 in every vector
    enc = fa.encode_from_memory(x, 27)
    bw0 = int(1023 * K).bit_length()
    dict_len = 8 + 12 + 128 * bw0 + 12
    assert enc[0] == 8 and enc[dict_len] == 11  # dictionary page, then the DICT_RLE page
    want = (100 + np.arange(1024) % 8).astype(ddt)
    packed = oracle.ffor(want, 3, ddt(100))
    hdr = np.zeros(12, dtype=np.uint8)
    hdr[0] = 3
    hdr[2:4] = np.frombuffer(np.uint16(2).tobytes(), np.uint8)
    hdr[4:12] = np.frombuffer(np.uint64(100).tobytes(), np.uint8)
    pos = np.array([7, 1000], dtype=np.uint16)
    exc = np.array([123456, 42], dtype=ddt)
    tail = np.zeros(12, dtype=np.uint8)
    tail[0] = 255
    blob = np.concatenate([enc[:8], hdr, packed.view(np.uint8), pos.view(np.uint8), exc.view(np.uint8), tail,
                           enc[dict_len:]])
    want[7], want[1000] = 123456, 42
    expect = np.tile(want.astype(dtype), 8)
    np.testing.assert_array_equal(oracle.decode_column(blob, dtype, x.size), expect)
    plan = fa.Plan(blob, dtype)
    d_in, d_out = dev(torch, blob), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, dtype, x.size), expect)


def test_c_api_round_trip_host_pointers(torch_cuda):
    """decode_to_memory (the reference's C entry point) on a DICT_RLE column in host memory."""
    x = wl.dict_runs(8 * 1024, np.int64, seed=11)
    enc = fa.encode_from_memory(x, 27)
    out = np.zeros_like(x)
    fa.decode_to_memory(enc.ctypes.data, out.ctypes.data, np.int64)
    np.testing.assert_array_equal(out, x)


def test_empty_column_and_encoded_size(torch_cuda):
    """The plan's encoded size is the encoder's output size, for an empty column and a multi-chunk one."""
    torch = torch_cuda
    for x in (np.zeros(0, np.int64), wl.dict_runs(40 * 1024, np.int64, seed=13)):
        enc = fa.encode_from_memory(x, 27)
        plan = fa.Plan(enc, np.int64)
        assert plan.n_values == x.size and plan.encoded_size == enc.size
        d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
        plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
        plan.check()
        np.testing.assert_array_equal(to_np(d_out, np.int64, x.size), x)
