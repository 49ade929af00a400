"""Launch-shape autotuner for the decode kernels.

For every kernel family (raw streams per element width and bit width; UNCOMPRESSED / FFOR / DELTA / RLE /
dictionary pages per element width) it times candidate launch shapes -- threads per block, non-temporal stores,
a cap on the grid (persistent blocks), direct stores for wide values, and for narrow RLE pages the vectors per
wavefront (one vector with one position lane per thread, or two with two) -- on the GPU and writes the fastest into a
tuning table (`fastlanes_amd/tuning.txt` by default) that the library loads on first use. The row-group launch (family 6) is
tuned on the mixed 64-column row-group: KiB of decoded values per tile and blocks per compute unit. Entries of
families that a run does not tune are kept from the existing table.

    python -m fastlanes_amd.autotune [--out PATH] [--log2n 26] [--families stream,pages,rowgroup]
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import time

import numpy as np

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

FAM_STREAM, FAM_UNC, FAM_FFOR, FAM_DELTA, FAM_RLE, FAM_DICT, FAM_ROWGROUP = range(7)
RG_TILE_KIB = (16, 32, 64, 128)
RG_CU_BLOCKS = (0, 8, 16, 32, 64, 128)  # 0 = resident blocks only (one long run per wave)
BLOCKS = (128, 256, 512)
NTS = (0, 1)
CU_BLOCKS = (0, 1, 2, 4, 8)  # grid cap in blocks per compute unit (0 = one block per group of vectors)


def _time(fn, stream, reps=5, warm=2):
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


def _candidates(n_cu, nts=NTS):
    for block in BLOCKS:
        for nt in nts:
            for k in CU_BLOCKS:
                yield block, nt, k * n_cu


def tune_family(lib, family, W, bw, fn, stream, n_cu, bytes_moved):
    best = None
    # integer pages of 32/64-bit values may also write without LDS staging (bit 1 of the store flags)
    nts = NTS + (2, 3) if (family in (FAM_UNC, FAM_FFOR, FAM_DELTA) and W >= 32) or \
        (family in (FAM_RLE, FAM_DICT) and W == 64) else NTS
    # RLE pages of 8/16-bit values: one vector per wave with one position lane per thread, or (bit 2) two vectors per
    # wave with two lanes per thread and LDS staging -- the vectors-per-wavefront choice of this family
    if family == FAM_RLE and W <= 16:
        nts = NTS + (4, 5)
    for block, nt, maxb in _candidates(n_cu, nts):
        fa.check(lib.fls_gpu_tune_set(family, W, bw, block, nt, maxb), "tune_set")
        t = _time(fn, stream)
        if best is None or t < best[0]:
            best = (t, block, nt, maxb)
    fa.check(lib.fls_gpu_tune_set(family, W, bw, best[1], best[2], best[3]), "tune_set")
    return best[1], best[2], best[3], bytes_moved / best[0] / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.txt"))
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--families", default="stream,pages")
    ap.add_argument("--widths", default="8,16,32,64")
    a = ap.parse_args()

    import torch

    lib = fa.lib()
    lib.fls_gpu_tune_set.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
    dev = torch.device("cuda:0")
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    stream = torch.cuda.current_stream()
    sp = int(stream.cuda_stream)
    n = 1 << a.log2n
    n_vec = n // 1024
    widths = [int(x) for x in a.widths.split(",")]
    lines = [f"# family width bw block nontemporal max_blocks   (measured GB/s)  {time.strftime('%Y-%m-%d')}",
             f"# device {torch.cuda.get_device_name(dev)} with {n_cu} compute units, 2^{a.log2n} values per case"]
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    fams = a.families.split(",")
    if "stream" in fams:
        for W in widths:
            out = torch.empty(n * W // 8 + 16, dtype=torch.uint8, device=dev)
            for bw in range(W + 1):
                packed = torch.randint(0, 256, (max(n_vec * 128 * bw, 16),), dtype=torch.uint8, device=dev,
                                       generator=g)
                bases = torch.zeros(n_vec * W // 8 + 16, dtype=torch.uint8, device=dev)

                def fn(packed=packed, bases=bases, W=W, bw=bw, out=out):
                    fa.unffor_stream(W, bw, packed.data_ptr(), bases.data_ptr(), out.data_ptr(), n_vec, sp)

                bytes_moved = n * bw / 8 + n * W / 8 + n_vec * W / 8
                blk, nt, mb, gbs = tune_family(lib, FAM_STREAM, W, bw, fn, stream, n_cu, bytes_moved)
                lines.append(f"{FAM_STREAM} {W} {bw} {blk} {nt} {mb}   # {gbs:.0f}")
                print(lines[-1], flush=True)
                del packed
            del out
    if "pages" in fams:
        rng = np.random.default_rng(3)
        for W in widths:
            dt = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[W]
            m = min(n, 1 << 25)
            cases = {
                FAM_UNC: (0, rng.integers(0, np.iinfo(dt).max, m, dtype=dt, endpoint=True)),
                FAM_FFOR: (1, rng.integers(0, 1 << min(W - 1, 7), m, dtype=dt)),
                FAM_DELTA: (11, np.cumsum(rng.integers(0, 100, m)).astype(dt)),
                FAM_RLE: (6, wl.geometric_runs(m, 16.0, seed=1, value_bits=min(W - 1, 20)).astype(dt)),
                FAM_DICT: (17, rng.integers(0, 1 << min(W - 1, 12), m).astype(dt)),
            }
            for fam, (eid, x) in cases.items():
                enc = fa.encode_from_memory(x, eid)
                plan = fa.Plan(enc, dt)
                d_in = torch.from_numpy(enc).to(dev)
                d_out = torch.empty(x.nbytes + 16, dtype=torch.uint8, device=dev)

                def fn(plan=plan, d_in=d_in, d_out=d_out, cap=x.nbytes):
                    plan.decode(d_in.data_ptr(), d_out.data_ptr(), cap, sp)

                blk, nt, mb, gbs = tune_family(lib, fam, W, 255, fn, stream, n_cu, enc.size + x.nbytes)
                lines.append(f"{fam} {W} 255 {blk} {nt} {mb}   # {gbs:.0f}")
                print(lines[-1], flush=True)
                del plan, d_in, d_out
    if "rowgroup" in fams:
        from fastlanes_amd import rowgroup_bench as rb

        vals, encs, offs, _, enc_total = rb.build()
        host = np.zeros(enc_total + 16, dtype=np.uint8)
        for o, e in zip(offs, encs):
            host[o: o + e.size] = e
        d_enc = torch.from_numpy(host).to(dev)
        del host
        moved = sum(e.size for e in encs) + sum(x.nbytes for x in vals)
        best, d_out = None, None
        for kib in RG_TILE_KIB:
            for per_cu in RG_CU_BLOCKS:
                fa.check(lib.fls_gpu_tune_set(FAM_ROWGROUP, 0, 255, kib, 1, per_cu), "tune_set")
                rg = fa.RowGroup(encs, [x.dtype for x in vals], offs)
                if d_out is None:
                    d_out = torch.empty(rg.total_bytes + 256, dtype=torch.uint8, device=dev)
                t = _time(lambda: rg.decode(d_enc.data_ptr(), d_out.data_ptr(), 0, sp), stream)
                print(f"# row-group tile {kib} KiB, {per_cu} blocks per CU: {t * 1e3:.3f} ms", flush=True)
                if best is None or t < best[0]:
                    best = (t, kib, per_cu)
                del rg
        fa.check(lib.fls_gpu_tune_set(FAM_ROWGROUP, 0, 255, best[1], 1, best[2]), "tune_set")
        lines.append(f"{FAM_ROWGROUP} 0 255 {best[1]} 1 {best[2]}   # {moved / best[0] / 1e9:.0f}")
        print(lines[-1], flush=True)
        del d_enc, d_out
    # keep the existing entries of families this run did not tune
    tuned = {tuple(int(v) for v in ln.split()[:3]) for ln in lines if ln and not ln.startswith("#")}
    if os.path.exists(a.out):
        with open(a.out) as f:
            for ln in f.read().splitlines():
                if ln.strip() and not ln.startswith("#") and tuple(int(v) for v in ln.split()[:3]) not in tuned:
                    lines.append(ln)
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
