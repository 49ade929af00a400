"""One-pass FSST decode (offsets and bytes from one launch per page, segment bases by decoupled look-back) against
the oracle, bit-exact: vocabulary strings, escape runs, empty and long strings, the reference's own FSST table,
repeated calls on one plan mixed with the two-step form, and the capacity error."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from tests.test_gpu_decode import _mixed_strings, dev, empty, torch_cuda  # noqa: F401

pytestmark = pytest.mark.gpu


def _fused(torch, plan, d_in, n, cap):
    d_off = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    d_bytes = empty(torch, cap)
    plan.decode_strings_fused(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), cap)
    plan.check()
    return d_off.cpu().numpy(), d_bytes.cpu().numpy()


def test_fused_vocabulary_strings(torch_cuda, oracle):
    torch = torch_cuda
    offs, data = wl.vocab_strings(40 * 1024, seed=11)
    enc = oracle.encode_str_column(offs, data)
    plan = fa.Plan(enc, "str")
    n = plan.n_values
    d_in = dev(torch, enc)
    for rep in range(3):  # new epoch per call; the status words of the previous call must not be taken as ready
        print(f"fsst fused vocabulary rep={rep}", flush=True)
        got_off, got = _fused(torch, plan, d_in, n, data.size)
        np.testing.assert_array_equal(got_off, offs.astype(np.int64))
        np.testing.assert_array_equal(got[: data.size], data)
    # the two-step form on the same plan afterwards
    d_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    plan.string_offsets(d_in.data_ptr(), d_off.data_ptr())
    np.testing.assert_array_equal(d_off.cpu().numpy(), offs.astype(np.int64))
    d_bytes = empty(torch, data.size)
    plan.decode_strings(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size)
    plan.check()
    np.testing.assert_array_equal(d_bytes.cpu().numpy()[: data.size], data)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fused_escapes_empty_and_long_strings(torch_cuda, oracle, seed):
    """Escape runs (0xFF bytes) across share boundaries, empty strings (equal start positions), and segments too
    large for the LDS image (per-string form)."""
    torch = torch_cuda
    print(f"fsst fused mixed seed={seed}", flush=True)
    offs, data = _mixed_strings(seed)
    enc = fa.encode_strings(offs.astype(np.uint32), data)
    n = offs.size - 1
    o2, d2 = oracle.decode_str_column(enc, n, data.size + 64)
    np.testing.assert_array_equal(d2, data)
    plan = fa.Plan(enc, "str")
    got_off, got = _fused(torch, plan, dev(torch, enc), n, data.size)
    np.testing.assert_array_equal(got_off, offs.astype(np.int64))
    np.testing.assert_array_equal(got[: data.size], data)


def _strs_from(rows):
    offs = np.zeros(len(rows) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(r) for r in rows])
    return offs, np.frombuffer(b"".join(rows), dtype=np.uint8).copy()


@pytest.mark.parametrize("shape", ["all_empty", "one_long", "escapes_only", "tiny"])
def test_fused_segment_shapes(torch_cuda, oracle, shape):
    """Segments whose shares are empty, a segment of one 30 KB string among empty ones (byte-parallel count, too
    large for LDS), strings of escaped bytes only (every second byte a literal), one-byte strings."""
    torch = torch_cuda
    g = np.random.default_rng(7)
    rows = []
    for v in range(4):
        for i in range(1024):
            if shape == "all_empty":
                rows.append(b"" if v != 2 or i % 100 else b"abc")
            elif shape == "one_long":
                rows.append(bytes(g.integers(97, 100, 30_000, dtype=np.uint8).tobytes()) if i == 700 else b"")
            elif shape == "escapes_only":
                rows.append(bytes(g.integers(128, 256, int(g.integers(0, 9)), dtype=np.uint8).tobytes()))
            else:
                rows.append(bytes([97 + i % 26]))
    offs, data = _strs_from(rows)
    enc = fa.encode_strings(offs.astype(np.uint32), data)
    n = offs.size - 1
    o2, d2 = oracle.decode_str_column(enc, n, data.size + 64)
    np.testing.assert_array_equal(d2, data)
    print(f"fsst fused shape={shape} bytes={data.size} encoded={enc.size}", flush=True)
    plan = fa.Plan(enc, "str")
    got_off, got = _fused(torch, plan, dev(torch, enc), n, max(data.size, 1))
    np.testing.assert_array_equal(got_off, offs.astype(np.int64))
    np.testing.assert_array_equal(got[: data.size], data)


def test_fused_capacity_error(torch_cuda, oracle):
    torch = torch_cuda
    offs, data = wl.vocab_strings(4 * 1024, seed=3)
    enc = oracle.encode_str_column(offs, data)
    plan = fa.Plan(enc, "str")
    n = plan.n_values
    d_in = dev(torch, enc)
    d_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    d_bytes = empty(torch, data.size)
    plan.decode_strings_fused(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size // 2)
    with pytest.raises(RuntimeError, match="capacity"):
        plan.check()
    np.testing.assert_array_equal(d_off.cpu().numpy(), offs.astype(np.int64))  # offsets are still whole


def test_fused_multi_page_column(torch_cuda):
    """A string column over the 128 MiB page limit (2^16 vocabulary strings repeated 96 times, ~182 MB encoded, two
    pages): each page is its own launch and its segments look back into the previous page's; the decode reproduces
    the input's offsets and bytes."""
    torch = torch_cuda
    o0, d0 = wl.vocab_strings(1 << 16, seed=13)
    reps = 96
    data = np.tile(d0, reps)
    offs = np.concatenate([[0], np.cumsum(np.tile(np.diff(o0).astype(np.int64), reps))])
    enc = fa.encode_strings(offs.astype(np.uint32), data)
    plan = fa.Plan(enc, "str")
    assert plan.n_pages >= 2
    n = plan.n_values
    print(f"fsst fused multi-page: {n} strings, {plan.n_pages} pages", flush=True)
    got_off, got = _fused(torch, plan, dev(torch, enc), n, data.size)
    np.testing.assert_array_equal(got_off, offs)
    np.testing.assert_array_equal(got[: data.size], data)


def test_fused_decode_in_a_graph(torch_cuda, oracle):
    """The fused decode captured in a HIP graph and replayed three times, with eager decodes of the same plan between
    the replays (they advance its ticket counter and epoch) and the outputs cleared before each replay: every replay
    decodes the column bit-exact, because each call's ticket base and epoch are taken on the device."""
    torch = torch_cuda
    offs, data = wl.vocab_strings(12 * 1024, seed=21)
    enc = oracle.encode_str_column(offs, data)
    plan = fa.Plan(enc, "str")
    n = plan.n_values
    d_in = dev(torch, enc)
    d_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    d_bytes = empty(torch, data.size)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # one eager call first, outside the capture
        plan.decode_strings_fused(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size, s.cuda_stream)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.decode_strings_fused(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size,
                                  torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        print(f"fsst graph replay {rep}", flush=True)
        d_off.zero_()
        d_bytes.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        plan.check()
        np.testing.assert_array_equal(d_off.cpu().numpy(), offs.astype(np.int64), err_msg=f"replay {rep}")
        np.testing.assert_array_equal(d_bytes.cpu().numpy()[: data.size], data, err_msg=f"replay {rep}")
        # eager work on the same plan between replays
        e_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
        plan.string_offsets(d_in.data_ptr(), e_off.data_ptr())
        e_bytes = empty(torch, data.size)
        plan.decode_strings_fused(d_in.data_ptr(), e_off.data_ptr(), e_bytes.data_ptr(), data.size)
        plan.check()
        np.testing.assert_array_equal(e_bytes.cpu().numpy()[: data.size], data)
