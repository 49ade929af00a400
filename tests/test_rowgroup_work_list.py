"""The persistent row-group launch's work list, checked on the CPU (no GPU): the host planner's tiles cover every
vector of every fused page exactly once, stay inside their pages and outputs and inside the LDS the launch gives
each wave; the kernel's claim loop over the atomic counter takes every tile once and ends on every wave; and an
emulation of the kernel's per-thread index arithmetic (tests/rowgroup_emulator.py) decodes every tile to the
oracle's values."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from tests import rowgroup_emulator as emu


def _mixed(seed=3, n_lo=1024, n_hi=1 << 15):
    """The mixed row-group of the GPU row-group test, plus dictionaries with plain and delta-coded codes and both
    run-length forms."""
    g = np.random.default_rng(seed)
    cols = []
    for W in (8, 16, 32, 64):
        for kind, eid in wl.ROWGROUP_EXPRS.items():
            n = int(g.integers(n_lo // 1024, n_hi // 1024 + 1)) * 1024
            cols.append((wl.rowgroup_column(kind, W, n, seed=int(g.integers(1 << 30))), eid))
        dt = emu.W2DT[W]
        keys = g.integers(0, np.iinfo(dt).max, 300, dtype=dt, endpoint=True)
        for eid in (16, 22, 29):  # DICT plain codes, DICT delta codes, DICT_FRQ FFOR codes
            cols.append((keys[g.integers(0, 300, 5 * 1024)], eid))
        cols.append((wl.geometric_runs(7 * 1024, 8.0, seed=W, value_bits=min(W, 20)).astype(dt), 7))  # plain RLE
    cols.append((np.arange(5120, dtype=np.uint32) * 7, 0))  # UNCOMPRESSED
    cols.append((np.arange(3072, dtype=np.int64) - 1500, 1))  # signed
    cols.append((np.arange(2048, dtype=np.int64) * 1000003, 11))
    return cols


@pytest.fixture(scope="module")
def planned():
    cols = _mixed()
    encs = [(fa.encode_from_memory(x, eid), x.dtype) for x, eid in cols]
    offs, pos = [], 0
    for e, _ in encs:
        offs.append(pos)
        pos = (pos + e.size + 255) // 256 * 256
    host, tiles, wave_lds, _ = emu.fa_plan(encs, offs)
    return cols, encs, offs, host, tiles, wave_lds


def test_tiles_cover_every_fused_vector_once(planned):
    cols, encs, offs, host, tiles, wave_lds = planned
    assert tiles.size > 0
    assert (tiles["nv"] >= 1).all() and (tiles["nv"] <= 64).all()
    seen = {}
    for t in tiles:
        key = int(t["page"])
        cover = seen.setdefault(key, np.zeros(int(t["vec_c"]), dtype=np.int32))
        cover[int(t["v0"]): int(t["v0"]) + int(t["nv"])] += 1
    for key, cover in seen.items():
        assert (cover == 1).all(), f"page at {key}: vectors covered {set(cover.tolist())} times"
    # output ranges of different tiles never overlap
    spans = sorted((int(t["out"]) + int(t["v0"]) * 128 * emu.kind_parts(int(t["kind"]))[1],
                    int(t["out"]) + (int(t["v0"]) + int(t["nv"])) * 128 * emu.kind_parts(int(t["kind"]))[1]) for t in tiles)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert a1 <= b0


def test_fused_kinds_and_fallback_columns(planned):
    cols, encs, offs, host, tiles, wave_lds = planned
    fused_pages = {int(t["page"]) for t in tiles}
    fused_cols = {c for c, o in enumerate(offs) if any(o <= p < o + encs[c][0].size for p in fused_pages)}
    for c, (x, eid) in enumerate(cols):
        # plain RLE positions (7) have no persistent kind and decode through their per-page kernel
        assert (c in fused_cols) == (eid != 7), (c, eid)
    kinds = {emu.kind_parts(int(k)) for k in np.unique(tiles["kind"])}
    for fam in ("ffor", "delta", "rle", "dict1", "dict2", "dict7", "unc"):
        assert any(f == fam for f, _ in kinds), fam


def test_tiles_fit_the_launch_lds(planned):
    """Every tile's staged ranges are the vectors' bytes, back to back: range 0 holds each vector's rows (128 * bw
    bytes; UNCOMPRESSED values and codes at their fixed size), ranges 1 and 2 the lane bases (128 bytes per vector)
    or the run values; and they fit the LDS slot of one wave with range 0's pad row."""
    cols, encs, offs, host, tiles, wave_lds = planned
    for t in tiles:
        fam, W = emu.kind_parts(int(t["kind"]))
        pg = emu._Page(host, t)
        nv, v0 = int(t["nv"]), int(t["v0"])
        if fam in ("ffor", "delta", "dict2", "dict7"):
            O = 1 if fam in ("delta", "dict7") else 0
            rows = sum(128 * int(pg.b[pg.off[O] + v]) for v in range(v0, v0 + nv))
        elif fam == "unc":
            rows = 128 * W * nv
        elif fam == "dict1":
            rows = 4096 * nv
        else:
            rows = sum(128 * int(pg.b[pg.off[2] + v]) for v in range(v0, v0 + nv))
        assert int(t["rlen"][0]) == rows, (fam, W, v0)
        if fam in ("delta", "dict7"):
            assert int(t["rlen"][1]) == 128 * nv and int(t["rsrc"][1]) == pg.off[0] + 128 * v0
        if fam == "rle":
            r1 = pg.ep(v0 + nv, 0) if v0 + nv < int(t["vec_c"]) else pg.off[1] - pg.off[0]
            assert int(t["rlen"][1]) == r1 - pg.ep(v0, 0) and int(t["rlen"][2]) == 128 * nv
        end = max(int(t["rlds"][r]) + (int(t["rlen"][r]) + 15) // 16 * 16 + (128 if r == 0 else 0)
                  for r in range(int(t["n_rng"])))
        assert end + 16 <= wave_lds, (fam, W, end, wave_lds)
        assert int(t["chunks"]) == sum((int(t["rlen"][r]) + 15) // 16 for r in range(3))


@pytest.mark.parametrize("n_waves", [1, 7, 512])
def test_wave_runs_take_every_tile_once_and_balance(planned, n_waves):
    """The launch-order list split into one run per wave holds every tile exactly once, and the runs' estimated
    work differs by at most one tile's (longest tile first to the least-loaded wave)."""
    cols, encs, offs = planned[:3]
    host, tiles, wave_lds, runs = emu.fa_plan(encs, offs, waves=n_waves)
    _, ref, _, _ = emu.fa_plan(encs, offs)
    assert runs[0] == 0 and runs[-1] == tiles.size and (np.diff(runs.astype(np.int64)) >= 0).all()
    key = lambda t: (int(t["page"]), int(t["v0"]))
    assert sorted(key(t) for r in emu.wave_runs(runs) for t in tiles[r]) == sorted(key(t) for t in ref)
    loads = [sum(emu.tile_cost(t) for t in tiles[r]) for r in emu.wave_runs(runs)]
    if tiles.size >= n_waves:
        assert max(loads) - min(loads) <= max(emu.tile_cost(t) for t in tiles)


def test_emulated_kernel_decodes_every_tile_like_the_oracle(planned, oracle):
    cols, encs, offs, host, tiles, wave_lds = planned
    total = 0
    starts = []
    for x, _ in cols:
        starts.append(total)
        total += (x.nbytes + 255) // 256 * 256
    out = np.zeros(total + 256, dtype=np.uint8)
    err = 0
    for t in tiles:
        d = None
        if emu.kind_parts(int(t["kind"]))[0].startswith("dict"):
            c = (int(t["dict"]) - 1) // 65536  # planner's placeholder: column * 65536 + page + 1
            enc = encs[c][0]
            n_dict = int.from_bytes(enc[4:8].tobytes(), "little")
            D = np.uint64 if cols[c][0].dtype.itemsize == 8 else np.uint32
            d, _ = oracle.hzl_decode(enc[8:].copy(), n_dict, D)
        err |= emu.emulate_tile(host, t, out, d)
    assert err == 0
    fused = {int(t["out"]) for t in tiles}
    checked = 0
    for c, (x, eid) in enumerate(cols):
        if eid == 7:
            continue
        want = oracle.decode_column(encs[c][0], x.dtype, x.size)
        np.testing.assert_array_equal(want, x)
        got = out[starts[c]: starts[c] + x.nbytes].view(x.dtype)
        np.testing.assert_array_equal(got, want, err_msg=f"column {c} (expression {eid}, {x.dtype})")
        checked += 1
    assert checked == len([1 for _, e in cols if e != 7]) and fused
