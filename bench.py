#!/usr/bin/env python3
"""Decode benchmark: FastLanes bit-unpack / FFOR bit-width sweep on one MI355X.

Workload (one step): for every element type uint8/16/32/64 and every bit width w in [0, W], decode 2^28 values
FFOR-packed (per-vector base) and 2^28 values plain-packed, all inputs resident in HBM. Values are U[0, 2^w), so
their packed bits are uniform random bytes; bases are U[0, 2^(W-w)] (bound clamped at w = 0). Output in the
reference's unffor/unpack order. A bounded sample of every case is checked bit-exact against the CPU oracle in the
CPU-baseline leg.

Prints one JSON line (metric, value in Gvalues/s, roofline of the dominant kernel, CPU baseline of the
reference's own compiled unffor on the host cores), with summaries of the cascade, row-group and wizard
configurations under "configs" (--no-configs skips them); --detail adds the per-family page sub-benchmarks.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gvalues/s decoded + % HBM peak; learned vs exhaustive wizard time @<=2% size"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28, help="values per (type, width) case")
    ap.add_argument("--widths", default="8,16,32,64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--detail", action="store_true", help="also run the per-family page sub-benchmarks (full output)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the cascade (2^26 values per column), row-group and wizard summaries")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/traffic.txt: `<kernel> <bytes>` lines)."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            return float(parts[1])
    return None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(widths, threads, samples, log2n=24, passes=10):
    """The CPU leg: (1) checks the sampled GPU outputs against the oracle's unffor; (2) times the reference's
    compiled generated::unffor (oracle/_ref) on the host cores: every (type, width) FFOR case at 2^log2n values, on
    one thread and on `threads` threads (started before the timed passes), median of `passes` passes per case.
    Returns (baseline, mismatching cases)."""
    from oracle.py_oracle import Oracle, Reference, reference_available

    orc = Oracle()
    bad = []
    for (W, w, plain, v0, pk, bs, got) in samples:
        want = orc.unffor_stream(pk, 4, w, bs, got.dtype)
        if not np.array_equal(got, want):
            bad.append((W, w, plain, v0))
    if not reference_available():
        return None, bad
    ref = Reference()
    rng = np.random.default_rng(0)
    n_vec = (1 << log2n) // 1024
    packed = rng.integers(0, 256, n_vec * 128 * 64 + 64, dtype=np.uint8)  # sliced per case: the widest needs all
    rates = {}
    for nt in (1, threads):
        total_vals, total_t = 0, 0.0
        for W in widths:
            dt = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[W]
            out = np.empty(n_vec * 1024, dtype=dt)
            for w in range(W + 1):
                bases = rng.integers(0, 2 ** min(W - w, 62) + 1, n_vec).astype(dt)
                t = ref.lib.ref_time_unffor(W, packed.ctypes.data, out.ctypes.data, n_vec, w, bases.ctypes.data, 0,
                                            passes, nt)
                total_vals += n_vec * 1024
                total_t += t
            del out
        rates[nt] = total_vals / total_t / 1e9
    return {"value": round(rates[threads], 3), "unit": "Gvalues/s", "cores": threads, "kind": "reference",
            "single_core": round(rates[1], 3), "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"reference generated::unffor, every (type,width) FFOR case of {widths} at 2^{log2n} values, "
                      f"median of {passes} passes per case after one untimed pass, on 1 thread (single_core) and "
                      f"on {threads} threads started before the timing"}, bad


def copy_bandwidth(dev, nbytes=1 << 30, reps=5):
    """The device's copy rate in GB/s (bytes read + written per second) for a contiguous nbytes copy: the box's
    attainable stream rate, reported beside the 8 TB/s peak."""
    import torch

    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / float(np.median(ts)) / 1e9


def other_configs(dev):
    """Summaries of the cascade decoders (2^26 values per column, 2^24 strings), the mixed 64-column row-group and the
    encoding wizard (512 columns of 2^20 values); each bit-exact-checked by its own module. A failure is reported
    in place of its summary."""
    import torch

    from fastlanes_amd import bench_detail, rowgroup_bench, wizard

    stream = torch.cuda.current_stream()
    out = {}
    keep = ("ms", "gvalues_s", "gstrings_s", "decoded_gb_s", "gb_s", "hbm_frac", "bit_exact")
    try:  # FFOR u32 bw 7, 2^20 values (1,000,000 + U[0, 2^7), seed 42), encoded by the library's own encoder
        import fastlanes_amd as fa
        from fastlanes_amd import workloads as wl

        x = wl.config1_values()
        enc = fa.encode_from_memory(x, 1)
        plan = fa.Plan(enc, x.dtype)
        d_in = torch.from_numpy(enc).to(dev)
        d_out = torch.empty(x.nbytes, dtype=torch.uint8, device=dev)
        sp = int(stream.cuda_stream)
        t = bench_detail._events_time(lambda: plan.decode(d_in.data_ptr(), d_out.data_ptr(), x.nbytes, sp), stream,
                                      reps=21, warm=5)
        ok = bool(np.array_equal(d_out.cpu().numpy().view(np.uint32), x))
        out["config1_ffor_u32_bw7"] = {"values": int(x.size), "us": round(t * 1e6, 2),
                                       "gvalues_s": round(x.size / t / 1e9, 2), "bit_exact": ok,
                                       "note": "5.1 MB of traffic: resident in the last-level cache, latency-bound"}
    except Exception as e:  # noqa: BLE001
        out["config1_ffor_u32_bw7"] = {"error": repr(e)}
    try:
        c = bench_detail.cascade(dev, stream, verbose=False)
        out["cascade"] = {k: {kk: v[kk] for kk in keep if kk in v} for k, v in c.items()}
    except Exception as e:  # noqa: BLE001 -- reported, not raised: the headline line must still print
        out["cascade"] = {"error": repr(e)}
    torch.cuda.empty_cache()
    try:
        r = rowgroup_bench.run(dev, stream)
        out["rowgroup"] = {k: r[k] for k in ("columns", "values", "decoded_bytes", "encoded_bytes", "graph", "branched",
                                             "per_page", "graph_vs_per_page", "persistent")}
    except Exception as e:  # noqa: BLE001
        out["rowgroup"] = {"error": repr(e)}
    torch.cuda.empty_cache()
    try:
        w = wizard.run(dev, stream)
        out["wizard"] = {k: w[k] for k in ("columns", "rowgroups", "exhaustive_ms", "learned_ms", "speedup",
                                           "test_regret", "test_accuracy", "no_network_test_regret",
                                           "no_network_test_accuracy", "train_s", "meets_targets")}
    except Exception as e:  # noqa: BLE001
        out["wizard"] = {"error": repr(e)}
    torch.cuda.empty_cache()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    import fastlanes_amd as fa

    fa.lib()
    dev = torch.device("cuda:0")
    widths = [int(x) for x in a.widths.split(",")]
    n = 1 << a.log2n
    n_vec = n // 1024
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    dt_of = {8: torch.uint8, 16: torch.int16, 32: torch.int32, 64: torch.int64}

    # ---- inputs resident in HBM --------------------------------------------------------------------------------
    cases = []  # (W, w, packed, bases)
    for W in widths:
        for w in range(W + 1):
            packed = torch.randint(0, 256, (max(n_vec * 128 * w, 16),), dtype=torch.uint8, device=dev, generator=g)
            hi = min(W - w, 62)
            bases = torch.randint(0, 2 ** hi + (1 if hi < 62 else 0), (n_vec,), dtype=torch.int64, device=dev,
                                  generator=g)
            bases = bases.to(dt_of[W]) if W < 64 else bases
            cases.append((W, w, packed, bases))
    outs = {W: torch.empty(n * W // 8 + 16, dtype=torch.uint8, device=dev) for W in widths}
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    sp = int(stream.cuda_stream)

    def launch(W, w, packed, bases, plain):
        fa.unffor_stream(W, w, packed.data_ptr(), 0 if plain else bases.data_ptr(), outs[W].data_ptr(), n_vec, sp)

    def step(events=None):
        for plain in (False, True):
            for (W, w, packed, bases) in cases:
                if events is not None:
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    launch(W, w, packed, bases, plain)
                    e1.record(stream)
                    events.append((W, w, plain, e0, e1))
                else:
                    launch(W, w, packed, bases, plain)

    # ---- bounded sample of every case's output (first and last 4 vectors), checked in the cpu_baseline leg ------
    samples = []
    for plain in (False, True):
        for (W, w, packed, bases) in cases:
            launch(W, w, packed, bases, plain)
            torch.cuda.synchronize()
            npdt = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[W]
            for v0 in (0, n_vec - 4):
                pk = packed[v0 * 128 * w:(v0 + 4) * 128 * w].cpu().numpy() if w else np.zeros(16, np.uint8)
                bs = np.zeros(4, npdt) if plain else bases[v0:v0 + 4].cpu().numpy().view(npdt)
                got = outs[W][v0 * 1024 * W // 8:(v0 + 4) * 1024 * W // 8].cpu().numpy().view(npdt)
                samples.append((W, w, plain, v0, pk, np.ascontiguousarray(bs), got))

    # ---- timed region --------------------------------------------------------------------------------------------
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(events)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per_case = {}
    for (W, w, plain, e0, e1) in events:
        per_case.setdefault((W, w, plain), []).append(e0.elapsed_time(e1) * 1e-3)
    values_per_step = 2 * len(cases) * n
    kernel_time = sum(sum(v) for v in per_case.values()) / a.steps

    def alg_bytes(W, w, plain):  # packed bits + stored values + (FFOR) one base per vector
        return n * w / 8 + n * W / 8 + (0 if plain else n_vec * W / 8)

    # dominant kernel: largest total time
    (dW, dw, dplain), dts = max(per_case.items(), key=lambda kv: sum(kv[1]))
    d_avg = sum(dts) / len(dts)
    d_bytes = alg_bytes(dW, dw, dplain)
    kname = f"k_unffor_stream<{dW},{dw}>"
    achieved = d_bytes / d_avg / 1e9
    traffic = traffic_from_profiles(kname)

    sweep = {}
    for W in widths:
        fr = []
        for w in range(W + 1):
            for plain in (False, True):
                t = sum(per_case[(W, w, plain)]) / len(per_case[(W, w, plain)])
                fr.append(alg_bytes(W, w, plain) / t / 1e9 / HBM_PEAK_GBS)
        sweep[f"u{W}"] = {"min_frac": round(min(fr), 3), "mean_frac": round(float(np.mean(fr)), 3),
                          "max_frac": round(max(fr), 3)}
    by_width = {}
    for W in widths:
        row = {"ffor": [], "unpack": []}
        for w in range(W + 1):
            for plain in (False, True):
                t = sum(per_case[(W, w, plain)]) / len(per_case[(W, w, plain)])
                row["unpack" if plain else "ffor"].append(round(alg_bytes(W, w, plain) / t / 1e9 / HBM_PEAK_GBS, 3))
        by_width[f"u{W}"] = row
    agg_bytes = sum(alg_bytes(W, w, p) for (W, w, p) in per_case)
    copy_gbs = copy_bandwidth(dev)
    u32w7 = per_case.get((32, 7, False))
    line = {
        "metric": METRIC,
        "value": round(values_per_step / (wall / a.steps) / 1e9, 2),
        "unit": "Gvalues/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u16/u32/u64",
        "data": "synthetic (seeded on-device random packed bytes and bases)",
        "config": {"workload": f"bit-width sweep: FFOR and plain unpack, uint8/16/32/64, every w in [0,W], "
                               f"2^{a.log2n} values per case", "cases": 2 * len(cases),
                   "values_per_step": values_per_step},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname + (" (plain unpack)" if dplain else " (FFOR)"),
                     "bytes_per_launch": d_bytes, "avg_launch_s": d_avg,
                     "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4)},
        "sweep_hbm_frac": sweep,
        "sweep_aggregate_hbm_frac": round(agg_bytes / kernel_time / 1e9 / HBM_PEAK_GBS, 4),
        "sweep_hbm_frac_by_bit_width": by_width,
        "u32_w7_ffor_hbm_frac": round(alg_bytes(32, 7, False) / (sum(u32w7) / len(u32w7)) / 1e9 / HBM_PEAK_GBS, 4)
        if u32w7 else None,
    }
    del cases, outs
    torch.cuda.empty_cache()
    if not a.no_configs:
        line["configs"] = other_configs(dev)
    if a.detail:
        from fastlanes_amd import bench_detail

        line["detail"] = {"pages": bench_detail.pages(dev, torch.cuda.current_stream(), verbose=False)}
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"], bad = cpu_baseline(widths, threads, samples)
        if bad:
            print(json.dumps({"error": "decode mismatch against the oracle", "cases": bad[:10]}))
            sys.exit(1)
        line["parity"] = f"bit-exact vs oracle on 8 sampled vectors of each of the {len(samples) // 2} cases"
    else:
        line["parity"] = "not checked (--no-cpu)"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
