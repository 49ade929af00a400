"""Sub-benchmarks behind `bench.py --detail`: page decode per expression family, the cascade configurations,
the row-group decode and the encoding wizard. Every figure is a median of timed launches with inputs resident in
HBM; bytes are the algorithmic bytes (encoded bytes read + decoded bytes written)."""
from __future__ import annotations

import time

import numpy as np

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

HBM = 8000.0


def _events_time(fn, stream, reps=5, warm=2):
    import torch

    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    return float(np.median(ts))


def decode_column_timed(x: np.ndarray, eid: int, dev, stream, check=True):
    import torch

    enc = fa.encode_from_memory(x, eid)
    plan = fa.Plan(enc, x.dtype)
    d_in = torch.from_numpy(enc).to(dev)
    d_out = torch.empty(x.nbytes + 16, dtype=torch.uint8, device=dev)
    sp = int(stream.cuda_stream)

    def fn():
        plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes, sp)

    t = _events_time(fn, stream)
    plan.check(sp)
    ok = None
    if check:
        got = d_out[: x.nbytes].cpu().numpy().view(x.dtype)
        ok = bool(np.array_equal(got, x))
    nbytes = enc.size + x.nbytes
    return {"values": int(x.size), "encoded_bytes": int(enc.size), "ms": round(t * 1e3, 4),
            "gvalues_s": round(x.size / t / 1e9, 2), "gb_s": round(nbytes / t / 1e9, 1),
            "hbm_frac": round(nbytes / t / 1e9 / HBM, 4), "bit_exact": ok, "pages": plan.n_pages}


def pages(dev, stream, log2n=26, verbose=True):
    """Every page family at each element width (bit-exact check against the input)."""
    n = 1 << log2n
    rng = np.random.default_rng(1)
    res = {}
    for W, dt in ((8, np.uint8), (16, np.uint16), (32, np.uint32), (64, np.uint64)):
        m = n if W < 64 else n // 2
        cases = {
            "uncompressed": (0, rng.integers(0, np.iinfo(dt).max, m, dtype=dt, endpoint=True)),
            "ffor_w7": (1, rng.integers(0, 1 << min(W - 1, 7), m, dtype=dt)),
            "delta": (11, np.cumsum(rng.integers(0, 100, m)).astype(dt)),
            "rle": (6, wl.geometric_runs(m, 16.0, seed=1, value_bits=min(W - 1, 20)).astype(dt)),
            "dict_ffor": (17, rng.integers(0, 1 << min(W - 1, 12), m).astype(dt)),
            "dict_rle": (27, wl.dict_runs(m, dt, n_keys=4096 if W > 8 else 200, mean=16.0, seed=2)),
        }
        if W >= 32:  # patched FFOR: a 7-bit body with 1% outliers anywhere in the type's range
            body = (1000 + rng.integers(0, 128, m)).astype(dt)
            hit = rng.random(m) < 0.01
            body[hit] = rng.integers(0, np.iinfo(dt).max, int(hit.sum()), dtype=dt, endpoint=True)
            for name, eid in (("ffor_s_patch", 2), ("ffor_b_patch", 4), ("ffor_ll_patch", 5)):
                cases[name] = (eid, body)
            if W == 64:
                cases["ffor_s_patch_rec"] = (3, body)
        for name, (eid, x) in cases.items():
            res[f"u{W}_{name}"] = decode_column_timed(x, eid, dev, stream)
            if verbose:
                print(f"page u{W} {name}: {res[f'u{W}_{name}']}", flush=True)
    return res


def cascade(dev, stream, log2n=26, log2_strings=24, verbose=True):
    """Configs: (a) int64 timestamps DELTA+FFOR, (b) uint32 runs RLE, (c) int32 Zipf dictionary codes, (d) FSST."""
    import torch

    n = 1 << log2n
    out = {}
    out["a_timestamps_delta"] = decode_column_timed(wl.timestamps(n), 11, dev, stream)
    out["b_runs_rle"] = decode_column_timed(wl.geometric_runs(n, 16.0, seed=4), 6, dev, stream)
    out["c_zipf_dict"] = decode_column_timed(wl.zipf_codes(n, 4096, 1.1, seed=5), 17, dev, stream)
    # (d) strings
    ns = 1 << log2_strings
    offs, data = wl.vocab_strings_fast(ns, seed=7)
    enc = fa.encode_strings(offs, data)
    plan = fa.Plan(enc, "str")
    d_in = torch.from_numpy(enc).to(dev)
    d_off = torch.empty(ns + 1, dtype=torch.int64, device=dev)
    d_bytes = torch.empty(data.size + 64, dtype=torch.uint8, device=dev)
    sp = int(stream.cuda_stream)

    def fn():  # offsets and bytes in one pass
        plan.decode_strings_fused(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size, sp)

    t = _events_time(fn, stream, reps=3)
    ok = bool(np.array_equal(d_bytes[: data.size].cpu().numpy(), data)) and bool(
        np.array_equal(d_off.cpu().numpy(), offs.astype(np.int64)))
    nbytes = enc.size + data.size + 8 * (ns + 1)
    out["d_strings_fsst"] = {"strings": ns, "raw_bytes": int(data.size), "encoded_bytes": int(enc.size),
                             "ms": round(t * 1e3, 3), "gstrings_s": round(ns / t / 1e9, 3),
                             "decoded_gb_s": round(data.size / t / 1e9, 1), "gb_s": round(nbytes / t / 1e9, 1),
                             "hbm_frac": round(nbytes / t / 1e9 / HBM, 4), "bit_exact": ok}
    if verbose:
        for k, v in out.items():
            print(f"cascade {k}: {v}", flush=True)
    return out


def select_bench(dev, stream, log2n=26, verbose=True):
    """Predicate pushdown against full decode on the same columns: `value < median` fused into the decode kernels
    (bitmap + counts written instead of values). Bytes are the algorithmic ones: encoded bytes read plus bitmap
    (1 bit per value) and counts written."""
    import torch

    n = 1 << log2n
    rng = np.random.default_rng(7)
    cases = {
        "u32_ffor_w7": (1, rng.integers(0, 128, n, dtype=np.uint32)),
        "i64_timestamps_delta": (11, wl.timestamps(n)),
        "u32_runs_rle": (6, wl.geometric_runs(n, 16.0, seed=1, value_bits=20).astype(np.uint32)),
        "i32_zipf_dict": (17, wl.zipf_codes(n, 4096, 1.1, seed=5, dtype=np.int32)),
        "i64_dict_runs_dict_rle": (27, wl.dict_runs(n, np.int64, n_keys=4096, mean=16.0, seed=2)),
    }
    out = {}
    sp = int(stream.cuda_stream)
    for name, (eid, x) in cases.items():
        enc = fa.encode_from_memory(x, eid)
        plan = fa.Plan(enc, x.dtype)
        d_in = torch.from_numpy(enc).to(dev)
        d_out = torch.empty(x.nbytes + 16, dtype=torch.uint8, device=dev)
        n_vec = n // 1024
        d_bm = torch.empty(16 * n_vec, dtype=torch.int64, device=dev)
        d_cnt = torch.empty(n_vec, dtype=torch.int32, device=dev)
        samp = np.sort(x[:: max(1, n // 4096)])
        c = samp[samp.size // 2]  # about half the values pass

        t_dec = _events_time(lambda: plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes, sp), stream)
        t_sel = _events_time(lambda: plan.select(d_in.data_ptr(), "<", c, d_bm.data_ptr(), d_cnt.data_ptr(), sp), stream)
        plan.check(sp)
        mask = x < c
        ok = bool(np.array_equal(d_cnt.cpu().numpy(), mask.reshape(-1, 1024).sum(axis=1)))
        sel_bytes = enc.size + n // 8 + 4 * n_vec
        out[name] = {"values": n, "decode_ms": round(t_dec * 1e3, 4), "select_ms": round(t_sel * 1e3, 4),
                     "select_gvalues_s": round(n / t_sel / 1e9, 1), "decode_gvalues_s": round(n / t_dec / 1e9, 1),
                     "select_hbm_frac": round(sel_bytes / t_sel / 1e9 / HBM, 4), "selected": int(mask.sum()),
                     "counts_exact": ok}
        if verbose:
            print(f"select {name}: {out[name]}", flush=True)
        del plan, d_in, d_out, d_bm, d_cnt
    return out


def encode_bench(dev, stream, log2n=26, verbose=True):
    """GPU encoder (fls_gpu_encode) against the host encoder (encode_from_memory, all host threads) on the same
    columns; bytes are the input values read plus the encoded bytes written."""
    import torch

    n = 1 << log2n
    rng = np.random.default_rng(9)
    cases = {"u32_ffor_w7": (1, rng.integers(0, 128, n, dtype=np.uint32)),
             "i64_timestamps_delta": (11, wl.timestamps(n)),
             "u64_uncompressed": (0, rng.integers(0, 1 << 62, n, dtype=np.uint64))}
    out = {}
    sp = int(stream.cuda_stream)
    for name, (eid, x) in cases.items():
        t0 = time.perf_counter()
        want = fa.encode_from_memory(x, eid)
        t_host = time.perf_counter() - t0
        d_x = torch.from_numpy(x).to(dev)
        d_out = torch.empty(want.size + 64, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize(dev)
        ts = []
        for _ in range(4):
            t0 = time.perf_counter()
            size = fa.gpu_encode(d_x.data_ptr(), n, x.dtype, eid, d_out.data_ptr(), want.size + 64, sp)
            ts.append(time.perf_counter() - t0)
        t = min(ts[1:])
        same = size == want.size and bool(torch.equal(d_out[:size].cpu(), torch.from_numpy(want)))
        out[name] = {"values": n, "encoded_bytes": int(size), "gpu_ms": round(t * 1e3, 3),
                     "gpu_gvalues_s": round(n / t / 1e9, 1), "gpu_gb_s": round((x.nbytes + size) / t / 1e9, 1),
                     "host_ms": round(t_host * 1e3, 1), "same_bytes": same}
        if verbose:
            print(f"encode {name}: {out[name]}", flush=True)
        del d_x, d_out
    return out


def spell_bench(dev, log2n=20, n_blocks=16, reps=5, verbose=True):
    """Structural wizard over a 64-column row-group of 2^20 rows (per block of four: a high-cardinality int64
    column, its copy, an int32 key with ~1000 distinct values and an int64 function of it), device-resident."""
    import torch

    from fastlanes_amd import spell as sp

    n = 1 << log2n
    g = np.random.default_rng(4)
    i = np.arange(n)
    cols, keep = [], []
    for blk in range(n_blocks):
        hi = g.integers(0, 1 << 62, n, dtype=np.int64)
        key = ((i * (blk + 3)) % (1000 + blk)).astype(np.int32)
        cols += [hi, hi.copy(), key, key.astype(np.int64) * 11 + blk]
    descs = []
    for c in cols:
        d = torch.from_numpy(c).to(dev)
        keep.append(d)
        descs.append((fa.dtype_code(c.dtype), d.data_ptr(), 0))
    torch.cuda.synchronize(dev)
    times = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        kinds, refs = sp.spell_device(descs, n)
        times.append(time.perf_counter() - t0)
    t = min(times[1:])
    nbytes = sum(c.nbytes for c in cols)
    out = {"columns": len(cols), "rows": n, "column_bytes": int(nbytes), "ms": round(t * 1e3, 3),
           "gb_s_over_column_bytes": round(nbytes / t / 1e9, 1),
           "plan": {k: kinds.count(k) for k in sp.KINDS}}
    if verbose:
        print(f"spell: {out}", flush=True)
    return out


def main():
    import argparse
    import json

    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--sections", default="pages,cascade", help="pages, cascade, rowgroup, wizard, spell, select, encode")
    ap.add_argument("--log2n", type=int, default=26)
    a = ap.parse_args()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    res = {}
    for sec in a.sections.split(","):
        if sec == "pages":
            res["pages"] = pages(dev, stream, a.log2n)
        elif sec == "cascade":
            res["cascade"] = cascade(dev, stream, a.log2n)
        elif sec == "rowgroup":
            from fastlanes_amd import rowgroup_bench

            res["rowgroup"] = rowgroup_bench.run(dev, stream)
        elif sec == "wizard":
            from fastlanes_amd import wizard

            res["wizard"] = wizard.bench(dev)
        elif sec == "spell":
            res["spell"] = spell_bench(dev)
        elif sec == "select":
            res["select"] = select_bench(dev, stream, a.log2n)
        elif sec == "encode":
            res["encode"] = encode_bench(dev, stream, a.log2n)
    print(json.dumps(res), flush=True)


def run_all(dev):
    import torch

    stream = torch.cuda.current_stream()
    res = {}
    t0 = time.time()
    res["pages"] = pages(dev, stream)
    res["cascade"] = cascade(dev, stream)
    try:
        from fastlanes_amd import rowgroup_bench

        res["rowgroup"] = rowgroup_bench.run(dev, stream)
    except ImportError:
        pass
    try:
        from fastlanes_amd import wizard

        res["wizard"] = wizard.bench(dev)
    except ImportError:
        pass
    res["seconds"] = round(time.time() - t0, 1)
    return res


if __name__ == "__main__":
    main()
