"""Row-group decode parity: many columns of mixed widths and encodings through fls_gpu_rowgroup_* -- the
persistent launch over the work list (mode 0), one launch per page (mode 1), the per-page launches over parallel
branches (mode 2), and the persistent launch captured once in a HIP graph and replayed -- against the CPU oracle,
bit-exact."""
import ctypes as C

import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

pytestmark = pytest.mark.gpu

W2DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def build(torch, oracle, columns):
    """columns: list of (values, expr_id). Returns (RowGroup, device encoded buffer, expected decodes)."""
    encs, dts, offs, want, pos = [], [], [], [], 0
    for x, eid in columns:
        enc = fa.encode_from_memory(x, eid)
        encs.append(enc)
        dts.append(x.dtype)
        offs.append(pos)
        pos = (pos + enc.size + 255) // 256 * 256
        exp = oracle.decode_column(enc, x.dtype, x.size)
        np.testing.assert_array_equal(exp, x)  # the oracle decodes the product encoder's bytes to the input
        want.append(exp)
    host = np.zeros(pos + 16, dtype=np.uint8)
    for o, e in zip(offs, encs):
        host[o: o + e.size] = e
    d_enc = torch.from_numpy(host).cuda()
    rg = fa.RowGroup(encs, dts, offs)
    return rg, d_enc, want


def verify(torch, rg, d_out, want, what):
    torch.cuda.synchronize()
    rg.check()
    host = d_out.cpu().numpy()
    for c, exp in enumerate(want):
        got = host[rg.offsets[c]: rg.offsets[c] + exp.nbytes].view(exp.dtype)
        np.testing.assert_array_equal(got, exp, err_msg=f"{what}: column {c} ({exp.dtype}, {exp.size} values)")


def mixed_columns(seed=3, n_lo=1024, n_hi=1 << 16):
    g = np.random.default_rng(seed)
    cols = []
    for W in (8, 16, 32, 64):
        for kind, eid in wl.ROWGROUP_EXPRS.items():
            n = int(g.integers(n_lo // 1024, n_hi // 1024 + 1)) * 1024  # columns hold whole vectors
            cols.append((wl.rowgroup_column(kind, W, n, seed=int(g.integers(1 << 30))), eid))
    cols.append((np.arange(5120, dtype=np.uint32) * 7, 0))  # UNCOMPRESSED
    cols.append((np.arange(3072, dtype=np.int64) - 1500, 1))  # signed
    return cols


def test_rowgroup_mixed_all_modes(torch_cuda, oracle):
    torch = torch_cuda
    print("rowgroup mixed: build", flush=True)
    rg, d_enc, want = build(torch, oracle, mixed_columns())
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    assert d_out.data_ptr() % 256 == 0
    for mode in (0, 1, 2):
        print(f"rowgroup mixed: mode {mode}", flush=True)
        d_out.fill_(0xA5)
        rg.decode(d_enc.data_ptr(), d_out.data_ptr(), mode, fa.stream_ptr())
        verify(torch, rg, d_out, want, f"mode {mode}")
    # graph: captured once, replayed with the program's own work in between, every replay checked
    for rep in range(3):
        print(f"rowgroup mixed: graph replay {rep}", flush=True)
        d_out.zero_()
        torch.cuda.synchronize()
        rg.decode_graph(d_enc.data_ptr(), d_out.data_ptr(), fa.stream_ptr())
        verify(torch, rg, d_out, want, f"graph replay {rep}")


def test_rowgroup_fallback_expressions(torch_cuda, oracle):
    """Expressions outside the persistent kernel's set (RLE with plain positions, DICT_RLE) decode through their
    per-page kernels in the same call; dictionaries with plain, FFOR or delta-coded codes and frequency-ordered
    dictionaries are tiles of the persistent launch."""
    torch = torch_cuda
    g = np.random.default_rng(9)
    cols = []
    for dt in (np.uint8, np.uint32, np.int64):
        cols.append((wl.geometric_runs(20480, 8.0, seed=2, value_bits=7).astype(dt), 7))
        keys = g.integers(0, 100, 40).astype(dt)
        for eid in (16, 22, 28, 29, 27):
            cols.append((keys[g.integers(0, 40, 17408)], eid))
        cols.append((wl.rowgroup_column("dict", 8 * np.dtype(dt).itemsize, 9216, 5).view(dt), 17))
    rg, d_enc, want = build(torch, oracle, cols)
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    print("rowgroup fallback: per-page launches, persistent launch and graph", flush=True)
    for mode in (1, 0):
        d_out.fill_(0x5A)
        rg.decode(d_enc.data_ptr(), d_out.data_ptr(), mode, fa.stream_ptr())
        verify(torch, rg, d_out, want, f"fallback mode {mode}")
    d_out.fill_(0x5A)
    rg.decode_graph(d_enc.data_ptr(), d_out.data_ptr(), fa.stream_ptr())
    verify(torch, rg, d_out, want, "fallback graph")


def test_rowgroup_multi_page_columns(torch_cuda, oracle):
    """Columns longer than one page's 65535 vectors split into several pages, each cut into tiles."""
    torch = torch_cuda
    n = 65536 * 1024 + 3 * 1024
    cols = [(wl.rowgroup_column("ffor", 8, n, 21), 1), (wl.rowgroup_column("delta", 16, n, 22), 11)]
    rg, d_enc, want = build(torch, oracle, cols)
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    print("rowgroup multi-page: persistent graph and per-page launches", flush=True)
    rg.decode_graph(d_enc.data_ptr(), d_out.data_ptr(), fa.stream_ptr())
    verify(torch, rg, d_out, want, "multi-page graph")
    d_out.zero_()
    rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 1, fa.stream_ptr())
    verify(torch, rg, d_out, want, "multi-page mode 1")


def test_rowgroup_spec_sample(torch_cuda, oracle):
    """A scaled-down draw of the 64-column mixed row-group (lengths 2^12..2^16 instead of 2^20..2^26)."""
    torch = torch_cuda
    cols = []
    for W, kind, n, seed in wl.rowgroup_spec(64, seed=17, min_log2=12, max_log2=16):
        cols.append((wl.rowgroup_column(kind, W, n, seed), wl.ROWGROUP_EXPRS[kind]))
    rg, d_enc, want = build(torch, oracle, cols)
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    print("rowgroup spec sample: per-page, persistent, graph", flush=True)
    for mode in (1, 0):
        d_out.zero_()
        rg.decode(d_enc.data_ptr(), d_out.data_ptr(), mode, fa.stream_ptr())
        verify(torch, rg, d_out, want, f"spec mode {mode}")
    d_out.zero_()
    rg.decode_graph(d_enc.data_ptr(), d_out.data_ptr(), fa.stream_ptr())
    verify(torch, rg, d_out, want, "spec graph")


def test_rowgroup_empty_and_errors(torch_cuda):
    torch = torch_cuda
    rg = fa.RowGroup([], [], [])
    assert rg.total_bytes == 0
    buf = torch.empty(256, dtype=torch.uint8, device="cuda")
    rg.decode(buf.data_ptr(), buf.data_ptr(), 1)
    rg.check()
    x = np.arange(4096, dtype=np.uint32)
    enc = fa.encode_from_memory(x, 1)
    rg = fa.RowGroup([enc], [np.uint32], [0])
    with pytest.raises(fa.FlsError):
        rg.decode(buf.data_ptr(), buf.data_ptr() + 4, 0)  # output not 256-byte aligned
    with pytest.raises(fa.FlsError):
        rg.decode(buf.data_ptr(), buf.data_ptr(), 3)  # unknown mode


def test_rowgroup_persistent_kernel(torch_cuda, oracle):
    """The persistent launch alone (every page of the mixed row-group is a tile of the work list), then replays of
    its graph with other work on the stream between them."""
    torch = torch_cuda
    cols = mixed_columns()
    rg, d_enc, want = build(torch, oracle, cols)
    tiles, blocks, wave_lds, runs = rg.work_list()
    print(f"rowgroup persistent: {tiles.size} tiles, {blocks} blocks, {wave_lds} B LDS per wave", flush=True)
    assert tiles.size > 0 and blocks > 0 and runs[-1] == tiles.size
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    d_out.fill_(0x3C)
    rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 0, fa.stream_ptr())
    verify(torch, rg, d_out, want, "mode 0")
    for rep in range(2):
        d_out.fill_(0xC3)
        scratch = torch.randn(1 << 20, device="cuda").sum()  # other work between the replays
        rg.decode_graph(d_enc.data_ptr(), d_out.data_ptr(), fa.stream_ptr())
        verify(torch, rg, d_out, want, f"persistent graph replay {rep}")
        del scratch


def test_rowgroup_persistent_grid_stride(torch_cuda, oracle):
    """Tiles of one vector and one block per compute unit: every wave decodes a long run of tiles, through every
    kind of the mixed row-group."""
    torch = torch_cuda
    lib = fa.lib()
    saved = [C.c_uint(0) for _ in range(3)]
    fa.check(lib.fls_gpu_tune_get(6, 0, 255, *[C.byref(x) for x in saved]), "tune_get")
    fa.check(lib.fls_gpu_tune_set(6, 0, 255, 1, 1, 1), "tune_set")  # 1 KiB tiles (one vector), one block per CU
    try:
        rg, d_enc, want = build(torch, oracle, mixed_columns(seed=5, n_hi=1 << 14))
    finally:
        fa.check(lib.fls_gpu_tune_set(6, 0, 255, *[x.value for x in saved]), "tune_set")
    tiles, blocks, _, runs = rg.work_list()
    assert (tiles["nv"] == 1).all() and (runs.size - 1) % blocks == 0 and runs[-1] == tiles.size
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    print(f"rowgroup persistent, one-vector tiles: {tiles.size} tiles over {blocks} blocks", flush=True)
    rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 0, fa.stream_ptr())
    verify(torch, rg, d_out, want, "one-vector tiles")


def test_rowgroup_large_single_page_columns(torch_cuda, oracle):
    """Long single-page columns (u64 DELTA of 2 * 10^7 values, u32 DELTA, u64 FFOR): many two-vector tiles per
    page, every vector checked; an earlier build of the kernel got one staged row wrong in about 0.3% of such
    vectors."""
    torch = torch_cuda
    cols = [(wl.rowgroup_column("delta", 64, 20000 * 1024, 17004), 11),
            (wl.rowgroup_column("delta", 32, 20000 * 1024, 17005), 11),
            (wl.rowgroup_column("ffor", 64, 12000 * 1024, 17006), 1)]
    rg, d_enc, want = build(torch, oracle, cols)
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device="cuda")
    tiles, blocks, _, _ = rg.work_list()
    print(f"rowgroup large single-page columns: {tiles.size} tiles over {blocks} blocks", flush=True)
    for rep in range(2):
        d_out.fill_(0x77)
        rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 0, fa.stream_ptr())
        verify(torch, rg, d_out, want, f"large columns, run {rep}")

