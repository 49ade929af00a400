"""Mixed row-group decode benchmark: 64 columns of uint8/16/32/64 with lengths uniform in [2^20, 2^26] values
(multiples of 1024) and encodings drawn from bit-packing (w in 1..32), FFOR, delta, RLE and dictionary. The same
device-resident row-group is decoded four ways -- the persistent launch over the row-group's work list, that launch
captured once in a HIP graph and replayed, one launch per page in column order (the per-column form), and those
launches spread over 4 parallel stream branches -- and every column is compared with its input values.

    python -m fastlanes_amd.rowgroup_bench [--min-log2 20] [--max-log2 26] [--columns 64] [--forms persistent,...]
                                           [--no-per-kind]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

HBM = 8000.0


def build(n_columns=64, seed=17, min_log2=20, max_log2=26):
    """Encode the row-group on the host. Returns (values, encoded, device offsets, spec)."""
    spec = wl.rowgroup_spec(n_columns, seed=seed, min_log2=min_log2, max_log2=max_log2)
    vals, encs, offs, pos = [], [], [], 0
    for W, kind, n, s in spec:
        x = wl.rowgroup_column(kind, W, n, s)
        e = fa.encode_from_memory(x, wl.ROWGROUP_EXPRS[kind])
        vals.append(x)
        encs.append(e)
        offs.append(pos)
        pos = (pos + e.size + 255) // 256 * 256
    return vals, encs, offs, spec, pos


def _time(fn, stream, reps, warm=2):
    import torch

    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    return float(np.median(ts))


ALL_FORMS = ("persistent", "graph", "branched", "per_page")


def run(dev, stream, n_columns=64, seed=17, min_log2=20, max_log2=26, reps=10, forms_run=ALL_FORMS, per_kind=True):
    import torch

    t0 = time.time()
    vals, encs, offs, spec, enc_total = build(n_columns, seed, min_log2, max_log2)
    t_build = time.time() - t0
    host = np.zeros(enc_total + 16, dtype=np.uint8)
    for o, e in zip(offs, encs):
        host[o: o + e.size] = e
    d_enc = torch.from_numpy(host).to(dev)
    del host
    rg = fa.RowGroup(encs, [x.dtype for x in vals], offs)
    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device=dev)
    sp = int(stream.cuda_stream)
    n_values = sum(x.size for x in vals)
    dec_bytes = sum(x.nbytes for x in vals)
    enc_bytes = sum(e.size for e in encs)
    algo = enc_bytes + dec_bytes

    def check(what):
        torch.cuda.synchronize()
        rg.check(sp)
        for c, x in enumerate(vals):
            got = d_out[rg.offsets[c]: rg.offsets[c] + x.nbytes].cpu().numpy().view(x.dtype)
            if not np.array_equal(got, x):
                raise AssertionError(f"row-group {what}: column {c} {spec[c]} differs")

    res = {"columns": n_columns, "values": n_values, "decoded_bytes": dec_bytes, "encoded_bytes": enc_bytes,
           "host_build_s": round(t_build, 1)}
    tiles, blocks, wave_lds, _ = rg.work_list()
    res["work_list"] = {"tiles": int(tiles.size), "blocks": blocks, "lds_per_wave": wave_lds}
    forms = {"persistent": lambda: rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 0, sp),
             "graph": lambda: rg.decode_graph(d_enc.data_ptr(), d_out.data_ptr(), sp),
             "branched": lambda: rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 2, sp),
             "per_page": lambda: rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 1, sp)}
    for name, fn in forms.items():
        if name not in forms_run:
            continue
        d_out.fill_(0)
        fn()
        check(name)
        t = _time(fn, stream, reps)
        gbs = algo / t / 1e9
        res[name] = {"ms": round(t * 1e3, 3), "gvalues_s": round(n_values / t / 1e9, 2), "gb_s": round(gbs, 1),
                     "hbm_frac": round(gbs / HBM, 4)}
    for name in ("graph", "persistent", "branched"):
        if name in res and "per_page" in res:
            res[f"{name}_vs_per_page"] = round(res["per_page"]["ms"] / res[name]["ms"], 3)
    if not per_kind:
        return res
    # where the per-column time goes: each column decoded alone, grouped by (encoding, width)
    per_kind = {}
    for c, (W, kind, n, _) in enumerate(spec):
        plan = fa.Plan(encs[c], vals[c].dtype)
        base = d_enc.data_ptr() + offs[c]
        dst = d_out.data_ptr() + rg.offsets[c]
        t = _time(lambda: plan.decode(base, dst, vals[c].nbytes, sp), stream, 3, warm=1)
        k = per_kind.setdefault(f"{kind}{W}", {"columns": 0, "ms": 0.0, "bytes": 0})
        k["columns"] += 1
        k["ms"] += t * 1e3
        k["bytes"] += vals[c].nbytes + encs[c].size
    # the same groups of columns as row-groups of their own, decoded by the persistent launch
    groups = {}
    for c, (W, kind, n, _) in enumerate(spec):
        groups.setdefault(f"{kind}{W}", []).append(c)
    for key, cs in groups.items():
        sub = fa.RowGroup([encs[c] for c in cs], [vals[c].dtype for c in cs], [offs[c] for c in cs])
        t = _time(lambda: sub.decode(d_enc.data_ptr(), d_out.data_ptr(), 0, sp), stream, 3, warm=1)
        per_kind[key]["persistent_ms"] = round(t * 1e3, 3)
        per_kind[key]["persistent_hbm_frac"] = round(per_kind[key]["bytes"] / t / 1e9 / HBM, 3)
        del sub
    for k in per_kind.values():
        k["hbm_frac"] = round(k["bytes"] / (k["ms"] * 1e-3) / 1e9 / HBM, 3)
        k["ms"] = round(k["ms"], 3)
    res["per_kind_alone"] = dict(sorted(per_kind.items(), key=lambda kv: -kv[1]["ms"]))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--min-log2", type=int, default=20)
    ap.add_argument("--max-log2", type=int, default=26)
    ap.add_argument("--columns", type=int, default=64)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--forms", default=",".join(ALL_FORMS), help="comma list of " + ", ".join(ALL_FORMS))
    ap.add_argument("--no-per-kind", action="store_true", help="skip the per-kind breakdown")
    a = ap.parse_args()
    import torch

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    print(json.dumps(run(dev, stream, a.columns, 17, a.min_log2, a.max_log2, a.reps, tuple(a.forms.split(",")),
                         not a.no_per_kind)), flush=True)


if __name__ == "__main__":
    main()
