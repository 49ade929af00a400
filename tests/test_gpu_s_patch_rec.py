"""S_PATCH_REC pages (pool id 3) on the GPU: decode and predicates through the C ABI against the CPU oracle."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from tests.test_gpu_decode import _bitmap_words, dev, empty, to_np, torch_cuda  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_vec", [4, 200])
@pytest.mark.parametrize("dtype", [np.int64, np.uint64])
def test_decode_matches_oracle(torch_cuda, oracle, dtype, n_vec):
    torch = torch_cuda
    print(f"s_patch_rec dtype={np.dtype(dtype).name} n_vec={n_vec}", flush=True)
    x = wl.patch_column(dtype, n_vec * 1024, seed=n_vec)
    enc = fa.encode_from_memory(x, 3)
    want = oracle.decode_column(enc, dtype, x.size)
    np.testing.assert_array_equal(want, x)
    plan = fa.Plan(enc, dtype)
    assert plan.encoded_size == enc.size
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, dtype, x.size), want)


def test_no_exceptions(torch_cuda):
    torch = torch_cuda
    x = (np.arange(8 * 1024) % 100).astype(np.int64)
    enc = fa.encode_from_memory(x, 3)
    plan = fa.Plan(enc, np.int64)
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, np.int64, x.size), x)


def test_predicates_match_numpy(torch_cuda, oracle):
    torch = torch_cuda
    x = wl.patch_column(np.int64, 48 * 1024, seed=5)
    enc = fa.encode_from_memory(x, 3)
    plan = fa.Plan(enc, np.int64)
    d_in = dev(torch, enc)
    n_vec = x.size // 1024
    for op, fn, c in [("<", np.less, np.int64(1060)), (">", np.greater, np.int64(5000)), ("==", np.equal, x[77])]:
        d_bm = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
        d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
        plan.select(d_in.data_ptr(), op, c, d_bm.data_ptr(), d_cnt.data_ptr())
        plan.check()
        mask = fn(x, c)
        np.testing.assert_array_equal(d_bm.cpu().numpy().view(np.uint64), _bitmap_words(mask), err_msg=op)
        np.testing.assert_array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1), err_msg=op)


def test_c_api_round_trip_host_pointers(torch_cuda):
    x = wl.patch_column(np.int64, 8 * 1024, seed=7)
    enc = fa.encode_from_memory(x, 3)
    out = np.zeros_like(x)
    fa.decode_to_memory(enc.ctypes.data, out.ctypes.data, np.int64)
    np.testing.assert_array_equal(out, x)
