"""fls_gpu_encode_noalloc (the encoder with its scratch in a fixed device array instead of stream allocations)
writes the same bytes as fls_gpu_encode, called through the C ABI: two calls in a row per expression, the second
on a smaller column (the array keeps the first call's contents), and inputs over its 2^24-value maximum are
rejected before any work."""
import ctypes as C

import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _column(dtype, eid, n, seed):
    g = np.random.default_rng(seed)
    info = np.iinfo(dtype)
    if eid in (6, 7):
        return wl.geometric_runs(n, 16.0, seed=seed, value_bits=min(20, info.bits - 1)).astype(dtype)
    if eid == 11:
        return np.cumsum(g.integers(0, 100, n)).astype(dtype)
    if eid == 1:
        return (g.integers(0, 1 << min(info.bits - 2, 13), n) + info.max // 4).astype(dtype)
    return g.integers(info.min, info.max, n, dtype=dtype, endpoint=True)


def _encode(torch, fn, x, eid, cap):
    """Encodes x with entry point fn (fls_gpu_encode or fls_gpu_encode_noalloc) into a buffer of `cap` bytes;
    returns the encoded bytes."""
    lib = fa.lib()
    d_x = torch.zeros(max(x.nbytes, 16), dtype=torch.uint8, device="cuda")
    if x.size:
        d_x[: x.nbytes] = torch.from_numpy(x.view(np.uint8)).cuda()
    d_out = torch.full((cap,), 0x5A, dtype=torch.uint8, device="cuda")
    size = C.c_uint64(0)
    rc = getattr(lib, fn)(d_x.data_ptr(), x.size, fa.dtype_code(x.dtype), eid, d_out.data_ptr(), cap, C.byref(size),
                          fa.stream_ptr())
    fa.check(rc, fn)
    return d_out[: size.value].cpu().numpy()


@pytest.mark.parametrize("eid", [0, 1, 11, 6, 7])
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32, np.int64])
def test_noalloc_encoder_matches_allocating_encoder(torch_cuda, dtype, eid):
    torch = torch_cuda
    for call, n in enumerate((300 * 1024, 37 * 1024)):  # the second call is smaller than the first
        print(f"encode_noalloc dtype={np.dtype(dtype).name} expr={eid} call={call} n={n}", flush=True)
        x = _column(dtype, eid, n, seed=17 * n + eid)
        host = fa.encode_from_memory(x, eid)
        want = _encode(torch, "fls_gpu_encode", x, eid, host.size + 4096)
        got = _encode(torch, "fls_gpu_encode_noalloc", x, eid, host.size + 4096)
        np.testing.assert_array_equal(got, want, err_msg=f"call {call}, n={n}")
        np.testing.assert_array_equal(want, host)


def test_noalloc_encoder_limits(torch_cuda):
    """Empty input encodes like fls_gpu_encode; more than 2^24 values is FLS_E_CAPACITY before any device work."""
    torch = torch_cuda
    x = np.zeros(0, np.uint32)
    print("encode_noalloc empty and over-limit", flush=True)
    np.testing.assert_array_equal(_encode(torch, "fls_gpu_encode_noalloc", x, 1, 4096),
                                  _encode(torch, "fls_gpu_encode", x, 1, 4096))
    lib = fa.lib()
    buf = torch.zeros(1 << 12, dtype=torch.uint8, device="cuda")
    size = C.c_uint64(0)
    rc = lib.fls_gpu_encode_noalloc(buf.data_ptr(), (1 << 24) + 1024, fa.dtype_code(np.uint32), 1, buf.data_ptr(),
                                    buf.numel(), C.byref(size), fa.stream_ptr())
    assert rc == -4  # FLS_E_CAPACITY
