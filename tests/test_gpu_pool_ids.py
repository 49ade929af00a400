"""GPU decode of the integer pool's chain, DELTA-patch and dictionary-patch forms (ids 8-10, 12-15, 18-21, 23-26,
30-38) against the CPU oracle, bit-exact: every id on every type it is defined for, the table behind the recorded
sizes, grid-capped launches (next-stride metadata prefetch), predicates and a larger column."""
import ctypes as C
import os

import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from tests.test_gpu_decode import PRED, _bitmap_words, dev, empty, to_np, torch_cuda  # noqa: F401
from tests.test_oracle_pages import NEW_IDS, _gen_5exp_table, _new_id_column

pytestmark = pytest.mark.gpu

DTYPES = [np.uint8, np.int16, np.uint32, np.int64, np.int8, np.uint16, np.int32, np.uint64]


def _decode(torch, oracle, x, eid, dtype):
    enc = oracle.encode_column(x, eid)
    want = oracle.decode_column(enc, dtype, x.size)
    np.testing.assert_array_equal(want, x)
    plan = fa.Plan(enc, dtype)
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    return to_np(d_out, dtype, x.size), want, plan, d_in


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("eid", sorted(NEW_IDS))
def test_new_pool_ids_match_oracle(torch_cuda, oracle, dtype, eid):
    bits = np.dtype(dtype).itemsize * 8
    if bits < NEW_IDS[eid]:
        pytest.skip("not defined for this width (patches: 32/64-bit values; u64 chains: 64-bit values)")
    print(f"pool id dtype={np.dtype(dtype).name} expr={eid}", flush=True)
    x = _new_id_column(dtype, eid, 40 * 1024, seed=eid + bits)
    got, want, _, _ = _decode(torch_cuda, oracle, x, eid, dtype)
    np.testing.assert_array_equal(got, want)


def test_recorded_table_every_id(torch_cuda, oracle):
    """The one-vector table whose encodings reproduce the recorded sizes: every integer pool id decodes on the GPU."""
    x = _gen_5exp_table()
    for eid in range(39):
        print(f"recorded table expr={eid}", flush=True)
        got, want, _, _ = _decode(torch_cuda, oracle, x, eid, np.int64)
        np.testing.assert_array_equal(got, want, err_msg=f"expr {eid}")


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_new_forms_grid_capped(torch_cuda, oracle, W):
    """The patched and chained forms with the grid capped to one or two blocks: each block strides over the vectors
    and takes the next-stride metadata prefetch (lane bases included). The default table is restored afterwards."""
    torch = torch_cuda
    lib = fa.lib()
    dt = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[W]
    eids = [e for e in (10, 12, 14, 15, 19, 20, 21, 24, 25, 26, 31, 36) if W >= NEW_IDS[e]]
    fam = {10: 4}  # RLE family; every patched form launches with the FFOR pages' shape
    try:
        for eid in eids:
            x = _new_id_column(dt, eid, 41 * 1024, seed=eid * 5 + W)
            enc = oracle.encode_column(x, eid)
            want = oracle.decode_column(enc, dt, x.size)
            plan = fa.Plan(enc, dt)
            d_in = dev(torch, enc)
            for block, maxb in ((64, 1), (256, 2)):
                print(f"grid-capped W={W} expr={eid} block={block} max_blocks={maxb}", flush=True)
                fa.check(lib.fls_gpu_tune_set(fam.get(eid, 2), W, 255, block, 1, maxb), "tune_set")
                d_out = empty(torch, x.nbytes)
                plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
                plan.check()
                np.testing.assert_array_equal(to_np(d_out, dt, x.size), want, err_msg=f"expr {eid} block {block}")
    finally:
        path = os.path.join(os.path.dirname(os.path.abspath(fa.__file__)), "tuning.txt")
        lib.fls_gpu_tune_load.argtypes = [C.c_char_p]
        assert lib.fls_gpu_tune_load(path.encode()) > 0


@pytest.mark.parametrize("eid", [9, 12, 14, 15, 19, 20, 25, 26, 36])
@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.int64])
def test_new_forms_predicates(torch_cuda, oracle, dtype, eid):
    """Fused decode + compare on the new forms against numpy comparisons of the oracle's decode."""
    bits = np.dtype(dtype).itemsize * 8
    if bits < NEW_IDS[eid]:
        pytest.skip("not defined for this width")
    torch = torch_cuda
    print(f"select dtype={np.dtype(dtype).name} expr={eid}", flush=True)
    n = 24 * 1024
    x = _new_id_column(dtype, eid, n, seed=eid + 2 * bits)
    enc = oracle.encode_column(x, eid)
    want_x = oracle.decode_column(enc, dtype, n)
    plan = fa.Plan(enc, dtype)
    d_in = dev(torch, enc)
    n_vec = n // 1024
    g = np.random.default_rng(eid)
    for op, fn in PRED.items():
        c = dtype(x[int(g.integers(0, n))])
        d_bm = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
        d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
        plan.select(d_in.data_ptr(), op, c, d_bm.data_ptr(), d_cnt.data_ptr())
        plan.check()
        mask = fn(want_x, c)
        np.testing.assert_array_equal(d_bm.cpu().numpy().view(np.uint64), _bitmap_words(mask), err_msg=f"{op} {c}")
        np.testing.assert_array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1), err_msg=op)


@pytest.mark.parametrize("eid", [13, 14, 15, 24, 37])
def test_new_forms_large_column(torch_cuda, oracle, eid):
    """2^20 int64 values (1024 vectors) of the patched DELTA and dictionary forms: many waves, long chains."""
    x = _new_id_column(np.int64, eid, 1 << 20, seed=eid)
    print(f"large expr={eid}", flush=True)
    got, want, _, _ = _decode(torch_cuda, oracle, x, eid, np.int64)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(fa.encode_from_memory(x, eid), oracle.encode_column(x, eid))


def test_delta_ll_chain_leaving_the_vector_is_reported(torch_cuda, oracle):
    """A DELTA LL_PATCH page whose link leaves the vector: the oracle refuses it and the GPU reports err bit 8
    (fls_gpu_plan_check fails) instead of writing outside the vector."""
    x = _gen_5exp_table()
    enc = oracle.encode_column(x, 14)
    off = np.frombuffer(enc[8:32].tobytes(), np.uint32)
    bad = enc.copy()
    bad[off[2] + 2: off[2] + 4] = np.frombuffer(np.uint16(1023).tobytes(), np.uint8)  # first position
    with pytest.raises(ValueError):
        oracle.decode_column(bad, np.int64, x.size)
    plan = fa.Plan(bad, np.int64)
    d_in, d_out = dev(torch_cuda, bad), empty(torch_cuda, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    with pytest.raises(fa.FlsError):
        plan.check()


def _dbl_runs_column(n_vec, seed):
    """float64 runs of 1-40 equal values over normal draws, with NaN payloads, both zeros and infinities mixed in."""
    g = np.random.default_rng(seed)
    n = n_vec * 1024
    vals = np.round(g.standard_normal(n), 1)
    vals[g.choice(n, n // 50, replace=False)] = -0.0
    vals[g.choice(n, n // 50, replace=False)] = 0.0
    vals[g.choice(n, n // 100, replace=False)] = np.inf
    nan = (np.uint64(0x7FF8000000000000) | g.integers(1, 1 << 30, n // 100).astype(np.uint64)).view(np.float64)
    vals[g.choice(n, n // 100, replace=False)] = nan
    return vals[np.repeat(np.arange(n), g.integers(1, 41, n))[:n]]


@pytest.mark.parametrize("eid", [6, 7, 8, 9, 10])
def test_double_rle_pages(torch_cuda, oracle, eid):
    """The double pool's RLE forms decode on the GPU to the oracle's bit patterns (runs split by IEEE !=: the run's
    first pattern for +0.0/-0.0 runs, each NaN its own run), with the grid capped too, and predicates on them compare
    as IEEE values (NaN selects nothing but !=, -0.0 equals 0.0)."""
    torch = torch_cuda
    lib = fa.lib()
    x = _dbl_runs_column(37, seed=eid)
    print(f"double RLE expr={eid}", flush=True)
    enc = oracle.encode_column(x, eid)
    want = oracle.decode_column(enc, np.float64, x.size)
    plan = fa.Plan(enc, np.float64)
    d_in = dev(torch, enc)
    d_out = empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, np.uint64, x.size), want.view(np.uint64))
    try:
        for block, maxb in ((64, 1), (256, 2)):
            print(f"double RLE expr={eid} block={block} max_blocks={maxb}", flush=True)
            fa.check(lib.fls_gpu_tune_set(4, 64, 255, block, 1, maxb), "tune_set")
            d_out = empty(torch, x.nbytes)
            plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
            plan.check()
            np.testing.assert_array_equal(to_np(d_out, np.uint64, x.size), want.view(np.uint64))
    finally:
        path = os.path.join(os.path.dirname(os.path.abspath(fa.__file__)), "tuning.txt")
        lib.fls_gpu_tune_load.argtypes = [C.c_char_p]
        assert lib.fls_gpu_tune_load(path.encode()) > 0
    n_vec = x.size // 1024
    for op, fn in PRED.items():
        for c in (0.0, -0.0, 0.5, np.inf, np.nan):
            print(f"double RLE expr={eid} select {op} {c}", flush=True)
            d_bm = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
            d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
            plan.select(d_in.data_ptr(), op, c, d_bm.data_ptr(), d_cnt.data_ptr())
            plan.check()
            with np.errstate(invalid="ignore"):
                mask = fn(want, c)
            np.testing.assert_array_equal(d_bm.cpu().numpy().view(np.uint64), _bitmap_words(mask), err_msg=f"{op} {c}")
            np.testing.assert_array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1))
