"""Launch one decode kernel (and a copy of known size) a few times, for hardware-counter runs.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -- python -m fastlanes_amd.profile_kernel --width 64 --bw 63

Prints the algorithmic bytes of one launch of each, so counter readings (KiB per dispatch) can be compared with
them: `copy_f32` reads and writes `copy_bytes` each, which calibrates FETCH_SIZE / WRITE_SIZE against bytes that
are known to cross HBM.
"""
from __future__ import annotations

import argparse
import json

import fastlanes_amd as fa


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=64)
    ap.add_argument("--bw", default="63", help="bit width, or a comma-separated list")
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--plain", action="store_true", help="plain unpack (no FFOR bases)")
    ap.add_argument("--copy-log2-bytes", type=int, default=30)
    ap.add_argument("--fsst-log2-strings", type=int, default=0,
                    help="instead: decode an FSST string column of 2^k vocabulary strings (offsets + bytes)")
    ap.add_argument("--fsst-two-step", action="store_true",
                    help="with --fsst-log2-strings: offsets, then bytes (two launches per page) instead of one pass")
    ap.add_argument("--dict-rle", action="store_true",
                    help="instead: decode a DICT_RLE column (pool id 27) of 2^log2n values of --width bits, runs of "
                         "mean length 16 over 4096 keys")
    ap.add_argument("--page-ffor", action="store_true",
                    help="instead: decode an FFOR column (pool id 1) of 2^log2n values of --width bits, values in "
                         "[0, 128) (bit width 7)")
    ap.add_argument("--page-kind", choices=("ffor", "bitpack", "delta", "rle", "dict", "runs16"),
                    help="instead: decode the row-group workload's column of that kind (pool id 1, 1, 11, 6, 17; runs16: uint32 geometric runs of mean 16, pool id 6) of "
                         "2^log2n values of --width bits")
    a = ap.parse_args()
    import torch

    if a.page_kind:
        import numpy as np

        from fastlanes_amd import workloads as wl

        eid = {"ffor": 1, "bitpack": 1, "delta": 11, "rle": 6, "dict": 17, "runs16": 6}[a.page_kind]
        if a.page_kind == "runs16":  # the cascade's RLE workload: uint32 geometric runs of mean length 16
            x = wl.geometric_runs(1 << a.log2n, 16.0, seed=4)
        else:
            x = wl.rowgroup_column(a.page_kind, a.width, 1 << a.log2n, seed=5,
                                   w=7 if a.page_kind in ("ffor", "bitpack") else None)
        enc = fa.encode_from_memory(x, eid)
        plan = fa.Plan(enc, x.dtype)
        d_in = torch.from_numpy(enc).cuda()
        d_out = torch.empty(x.nbytes + 16, dtype=torch.uint8, device="cuda")
        st = torch.cuda.current_stream()
        plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes, st.cuda_stream)  # untimed first pass
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(st)
        for _ in range(a.reps):
            plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes, st.cuda_stream)
        t1.record(st)
        torch.cuda.synchronize()
        plan.check()
        sec = t0.elapsed_time(t1) * 1e-3 / a.reps
        ok = bool(np.array_equal(d_out[: x.nbytes].cpu().numpy().view(x.dtype), x))
        print(json.dumps({"kind": a.page_kind, "values": x.size, "width": x.dtype.itemsize * 8, "encoded_bytes": int(enc.size),
                          "pages": plan.n_pages, "reps": a.reps, "bit_exact": ok, "ms": round(sec * 1e3, 4),
                          "hbm_frac": round((enc.size + x.nbytes) / sec / 8e12, 3)}), flush=True)
        return

    if a.page_ffor:
        import numpy as np

        dt = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[a.width]
        n = 1 << a.log2n
        x = np.random.default_rng(1).integers(0, 1 << min(a.width - 1, 7), n).astype(dt)
        enc = fa.encode_from_memory(x, 1)
        plan = fa.Plan(enc, dt)
        d_in = torch.from_numpy(enc).cuda()
        d_out = torch.empty(x.nbytes + 16, dtype=torch.uint8, device="cuda")
        for _ in range(a.reps):
            plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
        torch.cuda.synchronize()
        plan.check()
        ok = bool(np.array_equal(d_out[: x.nbytes].cpu().numpy().view(dt), x))
        print(json.dumps({"values": n, "width": a.width, "encoded_bytes": int(enc.size), "pages": plan.n_pages,
                          "reps": a.reps, "bit_exact": ok}), flush=True)
        return

    if a.dict_rle:
        import numpy as np

        from fastlanes_amd import workloads as wl

        dt = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[a.width]
        n = 1 << a.log2n
        x = wl.dict_runs(n, dt, n_keys=4096 if a.width > 8 else 200, mean=16.0, seed=2)
        enc = fa.encode_from_memory(x, 27)
        plan = fa.Plan(enc, dt)
        d_in = torch.from_numpy(enc).cuda()
        d_out = torch.empty(x.nbytes + 16, dtype=torch.uint8, device="cuda")
        for _ in range(a.reps):
            plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes)
        torch.cuda.synchronize()
        plan.check()
        ok = bool(np.array_equal(d_out[: x.nbytes].cpu().numpy().view(dt), x))
        print(json.dumps({"values": n, "width": a.width, "encoded_bytes": int(enc.size), "pages": plan.n_pages,
                          "reps": a.reps, "bit_exact": ok}), flush=True)
        return

    if a.fsst_log2_strings:
        import numpy as np

        from fastlanes_amd import workloads as wl

        ns = 1 << a.fsst_log2_strings
        offs, data = wl.vocab_strings_fast(ns, seed=7)
        enc = fa.encode_strings(offs, data)
        plan = fa.Plan(enc, "str")
        d_in = torch.from_numpy(enc).cuda()
        d_off = torch.empty(ns + 1, dtype=torch.int64, device="cuda")
        d_bytes = torch.empty(data.size + 64, dtype=torch.uint8, device="cuda")
        for _ in range(a.reps):
            if a.fsst_two_step:
                plan.string_offsets(d_in.data_ptr(), d_off.data_ptr())
                plan.decode_strings(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size)
            else:
                plan.decode_strings_fused(d_in.data_ptr(), d_off.data_ptr(), d_bytes.data_ptr(), data.size)
        torch.cuda.synchronize()
        plan.check()
        assert np.array_equal(d_bytes[: data.size].cpu().numpy(), data), "decoded bytes differ"
        assert np.array_equal(d_off.cpu().numpy(), offs.astype(np.int64)), "decoded offsets differ"
        print(json.dumps({"strings": ns, "raw_bytes": int(data.size), "encoded_bytes": int(enc.size),
                          "pages": plan.n_pages, "reps": a.reps}), flush=True)
        return

    dev = torch.device("cuda:0")
    n = 1 << a.log2n
    n_vec = n // 1024
    W = a.width
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    bws = [int(b) for b in a.bw.split(",")]
    bases = torch.zeros(n_vec * W // 8 + 16, dtype=torch.uint8, device=dev)
    out = torch.empty(n * W // 8 + 16, dtype=torch.uint8, device=dev)
    sp = fa.stream_ptr()
    for bw in bws:
        packed = torch.randint(0, 256, (max(n_vec * 128 * bw, 16),), dtype=torch.uint8, device=dev, generator=g)
        for _ in range(a.reps):
            fa.unffor_stream(W, bw, packed.data_ptr(), None if a.plain else bases.data_ptr(), out.data_ptr(), n_vec,
                             sp)
        del packed
    src = torch.ones(1 << (a.copy_log2_bytes - 2), dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(a.reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    for bw in bws:
        print(json.dumps({"kernel": f"k_unffor_stream<{W},{bw}>", "launch_order": bws.index(bw),
                          "algorithmic_bytes": n * bw / 8 + n * W / 8 + (0 if a.plain else n_vec * W / 8),
                          "read_bytes": n * bw / 8 + (0 if a.plain else n_vec * W / 8), "write_bytes": n * W / 8,
                          "reps": a.reps}), flush=True)
    print(json.dumps({"copy_kernel": "torch mul (float32, elementwise)", "copy_bytes": src.numel() * 4}), flush=True)


if __name__ == "__main__":
    main()
