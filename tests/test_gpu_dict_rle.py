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
