"""GPU decode parity: every kernel through the C ABI against the CPU oracle, bit-exact."""
import hashlib
import json
import os

import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

W2DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def dev(torch, arr: np.ndarray):
    """Device copy of a host byte/array buffer (16-byte padded)."""
    b = np.ascontiguousarray(arr).view(np.uint8).ravel()
    t = torch.zeros(b.size + 16, dtype=torch.uint8, device="cuda")
    if b.size:
        t[: b.size] = torch.from_numpy(b).cuda()
    return t


def empty(torch, nbytes):
    return torch.empty(max(nbytes, 16), dtype=torch.uint8, device="cuda")


def to_np(t, dtype, n):
    return t.cpu().numpy().view(np.uint8)[: n * np.dtype(dtype).itemsize].view(dtype)


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_unffor_stream_every_bit_width(torch_cuda, oracle, W):
    torch = torch_cuda
    dt = W2DT[W]
    n_vec = 24
    for bw in range(W + 1):
        print(f"stream W={W} bw={bw}", flush=True)
        vals, bases = wl.sweep_case(W, bw, n_vec, seed=1)
        x = (vals + np.repeat(bases, 1024)).astype(dt)
        packed = oracle.ffor_stream(x, bw, bases)
        want = oracle.unffor_stream(packed, n_vec, bw, bases, dt)
        np.testing.assert_array_equal(want, x)
        d_in, d_b, d_out = dev(torch, packed), dev(torch, bases), empty(torch, x.nbytes)
        fa.unffor_stream(W, bw, d_in.data_ptr(), d_b.data_ptr(), d_out.data_ptr(), n_vec)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(to_np(d_out, dt, x.size), want, err_msg=f"W={W} bw={bw}")
        # plain unpack (no bases)
        pk = oracle.ffor_stream(vals.astype(dt), bw, np.zeros(n_vec, dtype=dt))
        d_in = dev(torch, pk)
        fa.unffor_stream(W, bw, d_in.data_ptr(), None, d_out.data_ptr(), n_vec)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(to_np(d_out, dt, x.size), vals.astype(dt), err_msg=f"unpack W={W} bw={bw}")


def test_config1_golden(torch_cuda, oracle):
    torch = torch_cuda
    g = json.load(open(os.path.join(GOLDEN, "config1_golden.json")))
    x = wl.config1_values()
    n_vec = x.size // 1024
    bases = np.full(n_vec, g["base"], dtype=np.uint32)
    packed = oracle.ffor_stream(x, 7, bases)
    assert hashlib.sha256(packed.tobytes()).hexdigest() == g["packed_sha256"]  # same bytes as the reference encoder
    d_in, d_b, d_out = dev(torch, packed), dev(torch, bases), empty(torch, x.nbytes)
    fa.unffor_stream(32, 7, d_in.data_ptr(), d_b.data_ptr(), d_out.data_ptr(), n_vec)
    torch.cuda.synchronize()
    got = to_np(d_out, np.uint32, x.size)
    assert hashlib.sha256(got.tobytes()).hexdigest() == g["decoded_sha256"]


def _column(dtype, eid, n, seed):
    g = np.random.default_rng(seed)
    info = np.iinfo(dtype)
    if eid in (2, 4, 5):
        return wl.patch_column(dtype, n, seed=seed)
    if eid in (6, 7):
        return wl.geometric_runs(n, 16.0, seed=seed, value_bits=min(20, info.bits - 1)).astype(dtype)
    if eid >= 16:
        keys = g.integers(info.min, info.max, 4096 if info.bits > 8 else 200, dtype=dtype, endpoint=True)
        return keys[g.integers(0, keys.size, n)]
    if eid == 11:
        return np.cumsum(g.integers(0, 100, n)).astype(dtype)
    return g.integers(info.min, info.max, n, dtype=dtype, endpoint=True)


DTYPES = [np.uint8, np.int16, np.uint32, np.int64, np.int8, np.uint16, np.int32, np.uint64]


@pytest.mark.parametrize("eid", [0, 1, 6, 7, 11, 16, 17, 22, 28, 29])
@pytest.mark.parametrize("dtype", DTYPES)
def test_page_decode_matches_oracle(torch_cuda, oracle, dtype, eid):
    torch = torch_cuda
    print(f"page dtype={np.dtype(dtype).name} expr={eid}", flush=True)
    n = 40 * 1024
    x = _column(dtype, eid, n, seed=eid * 7 + np.dtype(dtype).itemsize)
    enc = oracle.encode_column(x, eid)
    want = oracle.decode_column(enc, dtype, n)
    np.testing.assert_array_equal(want, x)
    plan = fa.Plan(enc, dtype)
    assert plan.n_values == n
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, dtype, n), want)


@pytest.mark.parametrize("eid", [2, 4, 5])
@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.uint64, np.int64])
def test_patched_ffor_matches_oracle(torch_cuda, oracle, dtype, eid):
    """S_PATCH / B_PATCH / LL_PATCH pages from the oracle's encoder (the reference's exception choice), decoded on the
    GPU; the column mixes narrow bodies with outliers, constant, all-zero and full-range vectors."""
    torch = torch_cuda
    print(f"patched dtype={np.dtype(dtype).name} expr={eid}", flush=True)
    n = 40 * 1024
    x = wl.patch_column(dtype, n, seed=eid + np.dtype(dtype).itemsize)
    enc = oracle.encode_column(x, eid)
    want = oracle.decode_column(enc, dtype, n)
    np.testing.assert_array_equal(want, x)
    plan = fa.Plan(enc, dtype)
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, dtype, n), want)


@pytest.mark.parametrize("eid", [2, 4, 5])
def test_patched_ffor_large_round_trip(torch_cuda, eid):
    """2^20 uint64 values through this library's encoder and the GPU decoder (round trip; the CPU encoders agree
    byte for byte on smaller inputs)."""
    torch = torch_cuda
    x = wl.patch_column(np.uint64, 1 << 20, seed=40 + eid)
    enc = fa.encode_from_memory(x, eid)
    plan = fa.Plan(enc, np.uint64)
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    assert np.array_equal(to_np(d_out, np.uint64, x.size), x)


def test_ll_patch_chain_leaving_the_vector_is_reported(torch_cuda, oracle):
    """A linked list whose next hop leaves the vector: the kernel stops the chain and reports it."""
    torch = torch_cuda
    g = np.random.default_rng(3)
    t = (1000 + g.integers(0, 128, 1024)).astype(np.int64)
    t[[10, 400, 700, 1000, 1023]] = 41_000
    x = oracle.untranspose(t)
    enc = oracle.encode_column(x, 5)
    off = np.frombuffer(enc[8:28].tobytes(), np.uint32)
    bad = enc.copy()
    bad[off[1] + 2: off[1] + 4] = np.frombuffer(np.uint16(1023).tobytes(), np.uint8)  # first position
    plan = fa.Plan(bad, np.int64)
    d_in, d_out = dev(torch, bad), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    with pytest.raises(fa.FlsError):
        plan.check()


def test_multi_page_delta(torch_cuda, oracle):
    torch = torch_cuda
    x = wl.timestamps(65_600 * 1024)
    enc = oracle.encode_column(x, 11)
    plan = fa.Plan(enc, np.int64)
    assert plan.n_pages == 2
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    got = to_np(d_out, np.int64, x.size)
    assert np.array_equal(got, x)


def test_predicate_on_multi_page_column(torch_cuda, oracle):
    """Selections of a two-page column land at each page's own vector offsets."""
    torch = torch_cuda
    x = wl.timestamps(65_600 * 1024)
    enc = fa.encode_from_memory(x, 11)
    plan = fa.Plan(enc, np.int64)
    assert plan.n_pages == 2
    n_vec = x.size // 1024
    d_in = dev(torch, enc)
    d_bm = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
    d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
    c = x[x.size // 2 + 12345]
    plan.select(d_in.data_ptr(), ">=", c, d_bm.data_ptr(), d_cnt.data_ptr())
    plan.check()
    mask = x >= c
    np.testing.assert_array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1))
    np.testing.assert_array_equal(d_bm.cpu().numpy().view(np.uint64), _bitmap_words(mask))


def test_out_of_range_codes_are_clamped_and_reported(torch_cuda, oracle):
    torch = torch_cuda
    x = np.tile(np.arange(4, dtype=np.int64), 1024)
    enc = oracle.encode_column(x, 16)  # DICT, uncompressed codes
    # corrupt one code of vector 0 (codes buffer follows the dictionary page and the page metadata)
    page = enc.size - 4096
    bad = enc.copy()
    bad[page + 100: page + 104] = np.frombuffer(np.uint32(999).tobytes(), np.uint8)
    plan = fa.Plan(bad, np.int64)
    d_in, d_out = dev(torch, bad), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    with pytest.raises(fa.FlsError):
        plan.check()


@pytest.mark.parametrize("dtype", [np.uint8, np.int16])
def test_rle_positions_past_the_runs_are_clamped_and_reported(torch_cuda, oracle, dtype):
    """A narrow RLE page whose lane base points past its vector's runs: the one-lane-per-thread kernel takes its
    clamping path for that vector (err bit 1, fls_gpu_plan_check fails) and the other vectors decode exactly."""
    torch = torch_cuda
    x = wl.geometric_runs(8 * 1024, 8.0, seed=4, value_bits=7).astype(dtype)
    enc = oracle.encode_column(x, 6)
    off = np.frombuffer(enc[8:28].tobytes(), np.uint32)
    bad = enc.copy()
    lb = off[1] + 3 * 128 + 2 * 5  # vector 3 (fixed 128-byte lane bases), lane 5
    bad[lb: lb + 2] = np.frombuffer(np.uint16(2000).tobytes(), np.uint8)
    plan = fa.Plan(bad, dtype)
    d_in, d_out = dev(torch, bad), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    with pytest.raises(fa.FlsError):
        plan.check()
    got = to_np(d_out, dtype, x.size).reshape(8, 1024)
    keep = [v for v in range(8) if v != 3]
    np.testing.assert_array_equal(got[keep], x.reshape(8, 1024)[keep])


@pytest.mark.parametrize("W", [8, 16])
def test_rle_two_lane_kernel(torch_cuda, oracle, W):
    """Tuning bit 2 of the RLE family switches narrow pages back to the two-lanes-per-thread kernel (the one that
    also serves predicates and position bit widths above 1): both kernels against the oracle."""
    import ctypes as C

    torch = torch_cuda
    lib = fa.lib()
    dt = W2DT[W]
    x = _column(dt, 6, 41 * 1024, seed=W)
    enc = oracle.encode_column(x, 6)
    want = oracle.decode_column(enc, dt, x.size)
    plan = fa.Plan(enc, dt)
    d_in = dev(torch, enc)
    try:
        for nt in (1, 5):
            print(f"rle W={W} nt={nt}", flush=True)
            fa.check(lib.fls_gpu_tune_set(4, W, 255, 256, nt, 0), "tune_set")
            d_out = empty(torch, x.nbytes)
            plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
            plan.check()
            np.testing.assert_array_equal(to_np(d_out, dt, x.size), want, err_msg=f"nt {nt}")
    finally:
        path = os.path.join(os.path.dirname(os.path.abspath(fa.__file__)), "tuning.txt")
        lib.fls_gpu_tune_load.argtypes = [C.c_char_p]
        assert lib.fls_gpu_tune_load(path.encode()) > 0


def test_fsst_strings_match_oracle(torch_cuda, oracle):
    torch = torch_cuda
    offs, data = wl.vocab_strings(6 * 1024, seed=5)
    enc = oracle.encode_str_column(offs, data)
    plan = fa.Plan(enc, "str")
    n = plan.n_values
    assert n == offs.size - 1
    d_in = dev(torch, enc)
    d_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    plan.string_offsets(d_in.data_ptr(), d_off.data_ptr())
    got_off = d_off.cpu().numpy()
    np.testing.assert_array_equal(got_off, offs.astype(np.int64))
    total = int(got_off[-1])
    d_bytes = empty(torch, total)
    plan.decode_strings(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), total)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_bytes.cpu().numpy()[:total], data)


def _mixed_strings(seed):
    """Vectors of: vocabulary strings; strings of random bytes including runs of 0xFF (escape runs after
    compression); empty strings; 100-300-byte strings (their segments exceed the LDS buffers and decode from
    global memory)."""
    g = np.random.default_rng(seed)
    out = []
    words = wl.vocabulary(300, seed=seed)
    for v in range(8):
        for i in range(1024):
            kind = v % 4
            if kind == 0:
                out.append(b" ".join(words[int(k)] for k in g.integers(0, len(words), int(g.integers(1, 5)))))
            elif kind == 1:
                b = bytearray(g.integers(0, 256, int(g.integers(0, 40)), dtype=np.uint8).tobytes())
                if i % 3 == 0 and b:
                    r = int(g.integers(1, 6))
                    p = int(g.integers(0, len(b)))
                    b[p:p] = b"\xff" * r
                out.append(bytes(b))
            elif kind == 2:
                out.append(b"" if i % 2 else words[int(g.integers(0, len(words)))])
            else:
                out.append(bytes(g.integers(97, 123, int(g.integers(100, 300)), dtype=np.uint8).tobytes()))
    return wl._strs(out)


@pytest.mark.parametrize("seed", [1, 2])
def test_fsst_escapes_empty_and_long_strings(torch_cuda, oracle, seed):
    """Both decode forms (byte-parallel segments in LDS, and one string per thread from global memory) on strings
    with escape runs, empty strings and segments too large for LDS, encoded by this library's FSST encoder."""
    torch = torch_cuda
    print(f"fsst mixed strings seed={seed}", flush=True)
    offs, data = _mixed_strings(seed)
    enc = fa.encode_strings(offs.astype(np.uint32), data)
    n = offs.size - 1
    o2, d2 = oracle.decode_str_column(enc, n, data.size + 64)
    np.testing.assert_array_equal(d2, data)
    plan = fa.Plan(enc, "str")
    d_in = dev(torch, enc)
    d_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    plan.string_offsets(d_in.data_ptr(), d_off.data_ptr())
    np.testing.assert_array_equal(d_off.cpu().numpy(), offs.astype(np.int64))
    d_bytes = empty(torch, data.size)
    plan.decode_strings(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size)
    torch.cuda.synchronize()
    plan.check()
    np.testing.assert_array_equal(d_bytes.cpu().numpy()[: data.size], data)


def test_fsst_reference_table_golden(torch_cuda, oracle):
    """Strings compressed by the reference's own fsst_create/fsst_compress (fixture), packed into pages."""
    torch = torch_cuda
    g = np.load(os.path.join(GOLDEN, "fsst_golden.npz"))
    table, comp, coff, offs, data = g["table"], g["comp"], g["comp_offsets"], g["offsets"], g["data"]
    n = offs.size - 1
    # build one FSST page per the reference layout: [dict hdr][table] then page {bytes, transposed end offsets}
    pages = []
    vec_c = n // 1024
    ends_all, bytes_all, eps0, eps1 = [], [], [], []
    b0 = 0
    for v in range(vec_c):
        s0, s1 = coff[v * 1024], coff[(v + 1) * 1024]
        ends = (coff[v * 1024 + 1:(v + 1) * 1024 + 1] - s0).astype(np.uint32)
        ends_all.append(oracle.transpose(ends))
        bytes_all.append(comp[s0:s1])
        if v:
            eps0.append(b0)
            eps1.append(4096 * v)
        b0 += s1 - s0
    hdr = np.array([7, 1, 0, 0, vec_c & 255, vec_c >> 8, 0, 0], dtype=np.uint8)
    meta = 8 + 8 + 8 * (vec_c - 1)
    offs_arr = np.array([meta, meta + b0], dtype=np.uint32)
    eps = np.array([x for pair in zip(eps0, eps1) for x in pair], dtype=np.uint32)
    page = np.concatenate([hdr, offs_arr.view(np.uint8), eps.view(np.uint8), np.concatenate(bytes_all),
                           np.concatenate(ends_all).view(np.uint8)])
    dh = np.array([7, 0, 0, 0, 0, 0, 0, 0], dtype=np.uint8)
    enc = np.concatenate([dh, table, page])
    o2, d2 = oracle.decode_str_column(enc, n, data.size + 64)
    np.testing.assert_array_equal(d2, data)
    plan = fa.Plan(enc, "str")
    d_in = dev(torch, enc)
    d_off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    plan.string_offsets(d_in.data_ptr(), d_off.data_ptr())
    d_bytes = empty(torch, data.size)
    plan.decode_strings(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_off.cpu().numpy(), offs.astype(np.int64))
    np.testing.assert_array_equal(d_bytes.cpu().numpy()[: data.size], data)


@pytest.mark.parametrize("eid", [0, 1, 2, 4, 5, 11, 6, 16])
def test_c_api_round_trip_host_pointers(torch_cuda, eid):
    x = _column(np.int64, eid, 8 * 1024, seed=3)
    enc = fa.encode_from_memory(x, eid)
    out = np.zeros_like(x)
    fa.decode_to_memory(enc.ctypes.data, out.ctypes.data, np.int64)
    np.testing.assert_array_equal(out, x)


def test_c_api_device_pointers(torch_cuda):
    torch = torch_cuda
    x = _column(np.int64, 11, 8 * 1024, seed=4)
    enc = fa.encode_from_memory(x, 11)
    d_in, d_out = dev(torch, enc), empty(torch, x.nbytes)
    fa.decode_to_memory(d_in.data_ptr(), d_out.data_ptr(), np.int64)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(d_out, np.int64, x.size), x)


@pytest.mark.parametrize("eid", [1, 2, 4, 5, 11, 6, 7, 16, 17])
def test_random_access(torch_cuda, eid):
    import ctypes as C
    x = _column(np.int64, eid, 5 * 1024, seed=9)
    enc = fa.encode_from_memory(x, eid)
    lib = fa.lib()
    for v in (0, 3, 4):
        nb = C.c_uint64(0)
        pairs = (C.c_uint64 * 16)()
        lib.init_random_access(enc.ctypes.data, v, C.byref(nb), pairs)
        assert nb.value > 0
        bufs = [enc[pairs[2 * b]: pairs[2 * b] + pairs[2 * b + 1]].copy() for b in range(nb.value)]
        ptrs = (C.c_void_p * nb.value)(*[b.ctypes.data for b in bufs])
        out = np.zeros(1024, dtype=np.int64)
        lib.random_access(enc.ctypes.data, v, C.byref(nb), pairs, ptrs, out.ctypes.data)
        np.testing.assert_array_equal(out, x[v * 1024:(v + 1) * 1024], err_msg=f"expr {eid} vector {v}")


@pytest.mark.parametrize("dtype", [np.int64, np.uint32])
def test_dictionary_page_with_hzl_exceptions(torch_cuda, oracle, dtype):
    """A dictionary payload with exception chunks (as the reference's HZL compressor writes them,
    src/cor/eng/hzl_cmp.cpp; layout src/cor/eng/hzl_de_cmp.cpp:40-58) spliced into an encoded DICT+FFOR column."""
    torch = torch_cuda
    DW = 64 if np.dtype(dtype).itemsize == 8 else 32
    ddt = np.uint64 if DW == 64 else np.uint32
    K = 1000
    v = (np.arange(1024, dtype=np.uint64) * K + 5).astype(dtype)
    x = np.tile(v, 8)  # insertion-order codes: 0..1023 repeated
    enc = fa.encode_from_memory(x, 17)
    bw0 = int(1023 * K).bit_length()
    dict_len = 8 + 12 + 128 * bw0 + 12
    assert enc[0] == 8 and enc[dict_len + 0] == 8  # dictionary page, then the data page (EXP_DICT)
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
    torch.cuda.synchronize()
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, dtype, x.size), expect)


def _respace_ffor_page(enc: np.ndarray, W: int) -> np.ndarray:
    """Rewrite a single-page FFOR column so that its bw buffer holds 2 bytes per vector (bw, padding): a valid page
    whose metadata does not sit at fixed strides, which sends the kernels down the entry-point path."""
    hdr = enc[:8].copy()
    vec_c = int(hdr[4]) | (int(hdr[5]) << 8)
    arr_c = 3
    off = enc[8:8 + 4 * arr_c].view(np.uint32).astype(np.int64)
    ep = np.zeros((vec_c, arr_c), dtype=np.int64)
    if vec_c > 1:
        ep[1:] = enc[8 + 4 * arr_c: 8 + 4 * arr_c + 4 * arr_c * (vec_c - 1)].view(np.uint32).reshape(vec_c - 1, arr_c)
    sz = W // 8
    bws = np.array([enc[off[0] + ep[v, 0]] for v in range(vec_c)], dtype=np.uint8)
    bw_buf = np.zeros(2 * vec_c, dtype=np.uint8)
    bw_buf[0::2] = bws
    bw_buf[1::2] = 0xEE
    base_buf = enc[off[1]: off[1] + sz * vec_c]
    packed_buf = enc[off[2]: off[2] + int(ep[-1, 2]) + 128 * int(bws[-1])]
    new_ep = ep.copy()
    new_ep[:, 0] = 2 * np.arange(vec_c)
    meta = 8 + 4 * arr_c + 4 * arr_c * (vec_c - 1)
    o0 = meta
    o1 = o0 + bw_buf.size
    o2 = o1 + base_buf.size
    parts = [hdr, np.array([o0, o1, o2], dtype=np.uint32).view(np.uint8),
             new_ep[1:].astype(np.uint32).ravel().view(np.uint8), bw_buf, base_buf, packed_buf]
    return np.concatenate(parts)


@pytest.mark.parametrize("W", [8, 32, 64])
def test_ffor_page_with_strided_metadata(torch_cuda, oracle, W):
    torch = torch_cuda
    dt = W2DT[W]
    x = np.random.default_rng(W).integers(0, 1 << min(W - 1, 11), 40 * 1024).astype(dt)
    enc = fa.encode_from_memory(x, 1)
    page = _respace_ffor_page(enc, W)
    np.testing.assert_array_equal(oracle.decode_column(page, dt, x.size), x)
    plan = fa.Plan(page, dt)
    d_in, d_out = dev(torch, page), empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(d_out, dt, x.size), x)


@pytest.mark.parametrize("W", [32, 64])
def test_integer_pages_direct_stores(torch_cuda, oracle, W):
    """UNCOMPRESSED / FFOR / DELTA pages of 32/64-bit values (and RLE / dictionary pages of 64-bit values) with the
    launch shape that writes lane segments directly (no LDS staging), several grid shapes; the default table is
    restored afterwards."""
    import ctypes as C
    import os

    torch = torch_cuda
    lib = fa.lib()
    lib.fls_gpu_tune_set.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
    lib.fls_gpu_tune_load.argtypes = [C.c_char_p]
    dt = W2DT[W]
    try:
        fams = ((1, 0), (2, 1), (3, 11)) + (((4, 6), (4, 7), (5, 16), (5, 17), (5, 22)) if W == 64 else ())
        for fam, eid in fams:
            x = _column(dt, eid, 37 * 1024, seed=fam + W)
            enc = oracle.encode_column(x, eid)
            want = oracle.decode_column(enc, dt, x.size)
            plan = fa.Plan(enc, dt)
            d_in = dev(torch, enc)
            for block, nt, maxb in ((256, 2, 0), (128, 3, 0), (512, 2, 8)):
                print(f"direct W={W} family={fam} block={block} nt={nt} max_blocks={maxb}", flush=True)
                fa.check(lib.fls_gpu_tune_set(fam, W, 255, block, nt, maxb), "tune_set")
                d_out = empty(torch, x.nbytes)
                plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
                plan.check()
                np.testing.assert_array_equal(to_np(d_out, dt, x.size), want)
    finally:
        path = os.path.join(os.path.dirname(os.path.abspath(fa.__file__)), "tuning.txt")
        assert lib.fls_gpu_tune_load(path.encode()) > 0


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_page_families_grid_capped(torch_cuda, oracle, W):
    """Every page family at width W with the grid capped to one or two blocks, so each block strides over many
    vectors and takes the next-stride metadata prefetch (plain, RLE, dictionary and patched pages); against the
    oracle, bit-exact. The default table is restored afterwards."""
    import ctypes as C
    import os

    torch = torch_cuda
    lib = fa.lib()
    dt = W2DT[W]
    eids = (0, 1, 11, 6, 7, 16, 17, 22, 28, 29) + ((2, 4, 5) if W >= 32 else ())
    fam_of = {0: 1, 1: 2, 2: 2, 4: 2, 5: 2, 11: 3, 6: 4, 7: 4, 16: 5, 17: 5, 22: 5, 28: 5, 29: 5}
    try:
        for eid in eids:
            x = _column(dt, eid, 41 * 1024, seed=eid * 7 + W)
            enc = oracle.encode_column(x, eid)
            want = oracle.decode_column(enc, dt, x.size)
            np.testing.assert_array_equal(want, x)
            plan = fa.Plan(enc, dt)
            d_in = dev(torch, enc)
            for block, maxb in ((64, 1), (256, 2)):
                print(f"grid-capped W={W} expr={eid} block={block} max_blocks={maxb}", flush=True)
                fa.check(lib.fls_gpu_tune_set(fam_of[eid], W, 255, block, 1, maxb), "tune_set")
                d_out = empty(torch, x.nbytes)
                plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
                plan.check()
                np.testing.assert_array_equal(to_np(d_out, dt, x.size), want, err_msg=f"expr {eid} block {block}")
    finally:
        path = os.path.join(os.path.dirname(os.path.abspath(fa.__file__)), "tuning.txt")
        lib.fls_gpu_tune_load.argtypes = [C.c_char_p]
        assert lib.fls_gpu_tune_load(path.encode()) > 0


@pytest.mark.parametrize("eid", [16, 17, 22])
def test_zipf_int32_dictionary(torch_cuda, oracle, eid):
    """The categorical column of the cascade benchmark: int32 codes over 4096 keys drawn with Zipf s = 1.1 over key
    rank, as DICT with plain, FFOR and DELTA codes; 2^20 values, against the oracle."""
    torch = torch_cuda
    x = wl.zipf_codes(1 << 20, 4096, 1.1, seed=5, dtype=np.int32)
    enc = oracle.encode_column(x, eid)
    want = oracle.decode_column(enc, np.int32, x.size)
    np.testing.assert_array_equal(want, x)
    print(f"zipf int32 dictionary expr={eid}: {enc.size} B", flush=True)
    plan = fa.Plan(enc, np.int32)
    d_in = dev(torch, enc)
    d_out = empty(torch, x.nbytes)
    plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
    plan.check()
    np.testing.assert_array_equal(to_np(d_out, np.int32, x.size), want)
    np.testing.assert_array_equal(fa.encode_from_memory(x, eid), enc)


def _bitmap_words(mask: np.ndarray) -> np.ndarray:
    """SelectionDS bitmap words of a boolean mask: per 1024 values, 16 uint64 words, bit i of word j = value 64j+i."""
    bits = np.packbits(mask.reshape(-1, 64), axis=1, bitorder="little")
    return bits.view(np.uint64).reshape(-1)


PRED = {"==": np.equal, ">": np.greater, "<": np.less, ">=": np.greater_equal, "<=": np.less_equal,
        "!=": np.not_equal}


@pytest.mark.parametrize("eid", [0, 1, 2, 5, 6, 7, 11, 16, 17, 22, 29])
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32, np.int32, np.int64, np.uint64])
def test_predicate_pushdown_matches_numpy(torch_cuda, oracle, dtype, eid):
    """Fused decode + compare (every page kind) against numpy comparisons of the oracle's decode: bitmaps, counts,
    index arrays, and the AND of two selections."""
    if eid in (2, 5) and np.dtype(dtype).itemsize < 4:
        pytest.skip("patched FFOR exists for 32/64-bit values only")
    torch = torch_cuda
    print(f"select dtype={np.dtype(dtype).name} expr={eid}", flush=True)
    n = 24 * 1024
    x = _column(dtype, eid, n, seed=eid + 3 * np.dtype(dtype).itemsize)
    enc = oracle.encode_column(x, eid)
    want_x = oracle.decode_column(enc, dtype, n)
    plan = fa.Plan(enc, dtype)
    d_in = dev(torch, enc)
    n_vec = n // 1024
    g = np.random.default_rng(eid)
    consts = [x[int(g.integers(0, n))], x[int(g.integers(0, n))], np.iinfo(dtype).min, np.iinfo(dtype).max]
    bms = []
    for k, (op, fn) in enumerate(PRED.items()):
        c = dtype(consts[k % len(consts)])
        d_bm = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
        d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
        plan.select(d_in.data_ptr(), op, c, d_bm.data_ptr(), d_cnt.data_ptr())
        plan.check()
        mask = fn(want_x, c)
        got_bm = d_bm.cpu().numpy().view(np.uint64)
        np.testing.assert_array_equal(got_bm, _bitmap_words(mask), err_msg=f"{op} {c}")
        np.testing.assert_array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1), err_msg=op)
        bms.append((d_bm, mask))
    # index arrays (the reference's idx_arr form) and AND of two selections
    d_idx = torch.zeros(n_vec * 1024, dtype=torch.int16, device="cuda")
    d_cnt = torch.zeros(n_vec, dtype=torch.int32, device="cuda")
    fa.selection_indices(bms[1][0].data_ptr(), n_vec, d_idx.data_ptr(), d_cnt.data_ptr())
    idx = d_idx.cpu().numpy().view(np.uint16).reshape(n_vec, 1024)
    cnt = d_cnt.cpu().numpy()
    for v in range(n_vec):
        np.testing.assert_array_equal(idx[v, : cnt[v]], np.nonzero(bms[1][1][v * 1024:(v + 1) * 1024])[0])
    d_and = torch.zeros(16 * n_vec, dtype=torch.int64, device="cuda")
    fa.selection_and(bms[1][0].data_ptr(), bms[3][0].data_ptr(), n_vec, d_and.data_ptr(), d_cnt.data_ptr())
    both = bms[1][1] & bms[3][1]
    np.testing.assert_array_equal(d_and.cpu().numpy().view(np.uint64), _bitmap_words(both))
    np.testing.assert_array_equal(d_cnt.cpu().numpy(), both.reshape(-1, 1024).sum(axis=1))


def test_predicate_on_doubles(torch_cuda):
    """DOUBLE columns (UNCOMPRESSED) compare as IEEE values: NaN selects nothing but !=, -0.0 equals 0.0."""
    torch = torch_cuda
    g = np.random.default_rng(2)
    x = np.round(g.standard_normal(4096), 1)
    x[::97] = np.nan
    x[5::31] = -0.0
    x[7::29] = 0.0
    enc = fa.encode_from_memory(x, 0)
    plan = fa.Plan(enc, np.float64)
    d_in = dev(torch, enc)
    for op, fn in PRED.items():
        for c in (0.0, -0.0, 0.5, np.nan):
            d_bm = torch.zeros(64, dtype=torch.int64, device="cuda")
            d_cnt = torch.zeros(4, dtype=torch.int32, device="cuda")
            plan.select(d_in.data_ptr(), op, c, d_bm.data_ptr(), d_cnt.data_ptr())
            with np.errstate(invalid="ignore"):
                mask = fn(x, c)
            np.testing.assert_array_equal(d_bm.cpu().numpy().view(np.uint64), _bitmap_words(mask), err_msg=f"{op} {c}")


@pytest.mark.parametrize("eid", [0, 1, 11, 6, 7])
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32, np.int32, np.int64, np.uint64])
def test_gpu_encoder_matches_host_encoder(torch_cuda, dtype, eid):
    """fls_gpu_encode writes the same bytes as the host encoder (itself pinned to the oracle and the reference's
    recorded sizes), and the result decodes back to the input."""
    torch = torch_cuda
    print(f"gpu encode dtype={np.dtype(dtype).name} expr={eid}", flush=True)
    for n in (0, 1024, 37 * 1024):
        x = _column(dtype, eid, n, seed=n + eid) if n else np.zeros(0, dtype)
        want = fa.encode_from_memory(x, eid)
        d_x = dev(torch, x)
        cap = want.size + 4096
        d_out = torch.zeros(cap, dtype=torch.uint8, device="cuda")
        size = fa.gpu_encode(d_x.data_ptr(), n, dtype, eid, d_out.data_ptr(), cap)
        assert size == want.size
        np.testing.assert_array_equal(d_out[:size].cpu().numpy(), want, err_msg=f"n={n}")


def test_gpu_encoder_multi_page_and_capacity(torch_cuda):
    """A column over 65535 vectors (two pages) and an 8-byte column whose pages hit the 128 MiB limit first; a too
    small capacity is reported with the needed size."""
    torch = torch_cuda
    x = wl.timestamps(65_600 * 1024)
    x[: 50_000 * 1024] //= 1000  # runs in the first part, for the RLE pages
    for eid in (11, 0, 6):
        want = fa.encode_from_memory(x, eid)
        d_x = dev(torch, x)
        d_out = torch.zeros(want.size + 64, dtype=torch.uint8, device="cuda")
        size = fa.gpu_encode(d_x.data_ptr(), x.size, np.int64, eid, d_out.data_ptr(), want.size + 64)
        assert size == want.size
        assert torch.equal(d_out[:size].cpu(), torch.from_numpy(want)), eid
        del d_out
    with pytest.raises(fa.FlsError):
        fa.gpu_encode(d_x.data_ptr(), x.size, np.int64, 11, d_x.data_ptr(), 100)
