"""CPU emulation of the persistent row-group kernel (fastlanes_amd/csrc/k_rowgroup.hip), test infrastructure only.

It replays what the kernel does on a tile list with numpy, one wave of 64 threads at a time: the claim loop over
the atomic counter, and every tile's per-thread index arithmetic as the kernel writes it (which staged row and bit
offset each value comes from, the split-lane DELTA prefix, the store offset of every value). The tests compare its
output with the oracle's decode, so a wrong lane, bit offset, prefix or store position in the kernel's mapping
shows up on the CPU before the kernel runs on a GPU.
"""
from __future__ import annotations

import numpy as np

import fastlanes_amd as fa

ORDER8 = [0, 4, 2, 6, 1, 5, 3, 7]
W2DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def order8(b):
    return np.asarray(ORDER8)[b]


def seg_start(W, i):
    i = np.asarray(i)
    if W in (8, 16):
        return 64 * (i % 16) + 8 * order8(i // 16)
    if W == 32:
        return 64 * (i % 16) + 32 * (i // 16)
    return 64 * i


def j_of_p16(p):
    return 2 * (p % 8) + p // 8


def kind_parts(kind):
    """(family, W) of an RgKind: family in unc / ffor / delta / rle / dict1 / dict2 / dict7."""
    fam = ("unc", "ffor", "delta", "rle", "dict1", "dict2", "dict7")[kind // 4]
    return fam, 8 << (kind % 4)


def wave_runs(runs):
    """The kernel's dispatch: wave w loops over tiles [runs[w], runs[w + 1]) of the launch-order list; returns the
    tile ids each wave decodes."""
    return [list(range(int(runs[w]), int(runs[w + 1]))) for w in range(len(runs) - 1)]


def tile_cost(t):
    """The host's estimate of a tile's work (rg_assign_waves): decoded and staged bytes plus fixed costs."""
    return int(t["nv"]) * (128 * kind_parts(int(t["kind"]))[1] + 2048) + 16 * int(t["chunks"]) + 4096


class _Page:
    def __init__(self, enc, t):
        self.b = enc[int(t["page"]):]
        self.arr_c = int(t["arr_c"])
        self.off = [int(x) for x in t["off"]]

    def u(self, at, n):
        return int.from_bytes(self.b[at: at + n].tobytes(), "little")

    def ep(self, v, buf):
        return 0 if v == 0 else self.u(8 + 4 * self.arr_c + 4 * ((v - 1) * self.arr_c + buf), 4)


def _lane_value(rows, i, bit, bw, W):
    """The kernel's lane_value<W>: bits [s, s + bw) of row w + 1 : row w of lane i, w = bit / W, s = bit % W."""
    w, s = bit // W, bit % W
    lo = rows[w, i].astype(object)
    hi = rows[w + 1, i].astype(object)
    x = ((hi << W) | lo) >> s.astype(object)
    mask = (1 << bw) - 1
    return (x & mask).astype(np.uint64)


def _wrap(x, W):
    return (np.asarray(x, dtype=object) % (1 << W)).astype(np.uint64)


def _int_vector(rows, bw, base, lb, W, delta):
    """Values of one vector as the kernel's threads produce and store them: returns out[1024] (untransposed)."""
    out = np.zeros(1024, dtype=np.uint64)
    t = np.arange(64)
    r = np.arange(16)
    if W in (64, 32):
        LW = 16 if W == 64 else 32
        i, g = t % LW, t // LW
        if W == 64:
            c = 8 * (r % 8) + 4 * (r // 8)
            gb = (2 * (g & 1) + (g >> 1)) * bw
            pos = (64 * i + 16 * g)[:, None] + r[None, :]
        else:
            c = 4 * (r % 8) + 2 * (r // 8)
            gb = g * bw
            pos = (seg_start(32, i) + 16 * g)[:, None] + r[None, :]
        bit = c[None, :] * bw + gb[:, None]
        v = _wrap(_lane_value(rows, np.broadcast_to(i[:, None], bit.shape), bit, bw, W).astype(object) + base, W)
        if delta:
            part = _wrap(np.cumsum(v.astype(object), axis=1), W)  # running sums of each thread
            tot = part[:, -1].astype(object)
            excl = np.zeros(64, dtype=object)
            for k in range(64):  # exclusive prefix over the thread's lane group (g = 0 .. g - 1)
                excl[k] = sum(tot[k - LW * m] for m in range(1, g[k] + 1))
            v = _wrap(part.astype(object) + (np.asarray(lb, dtype=object)[i] + excl)[:, None], W)
        out[pos.ravel()] = v.ravel()
    elif W == 16:
        j = j_of_p16(r)
        bit = np.broadcast_to(j[None, :] * bw, (64, 16))
        v = _wrap(_lane_value(rows, np.broadcast_to(t[:, None], bit.shape), bit, bw, W).astype(object) + base, W)
        if delta:
            v = _wrap(np.cumsum(v.astype(object), axis=1) + np.asarray(lb, dtype=object)[t][:, None], W)
        out[(seg_start(16, t)[:, None] + r[None, :]).ravel()] = v.ravel()
    else:
        r8 = np.arange(8)
        for lanes, shift in ((t, 0), (t + 64, 8)):
            bit = np.broadcast_to(r8[None, :] * bw, (64, 8))
            v = _wrap(_lane_value(rows, np.broadcast_to(lanes[:, None], bit.shape), bit, bw, W).astype(object) + base, W)
            if delta:
                v = _wrap(np.cumsum(v.astype(object), axis=1) + np.asarray(lb, dtype=object)[lanes][:, None], W)
            out[(seg_start(8, t)[:, None] + shift + r8[None, :]).ravel()] = v.ravel()
    return out


def stage_image(enc, t):
    """The wave's LDS slot after the kernel stages tile t: each range's bytes (whole 16-byte chunks) at its LDS
    offset; the pad row after range 0 and the rest of the slot hold garbage the decode must never use."""
    rng = np.random.default_rng(int(t["page"]) + int(t["v0"]))
    size = max(int(t["rlds"][r]) + (int(t["rlen"][r]) + 15) // 16 * 16 + 128 for r in range(int(t["n_rng"])))
    img = rng.integers(0, 256, size, dtype=np.uint8)
    page = enc[int(t["page"]):]
    for r in range(int(t["n_rng"])):
        n16 = (int(t["rlen"][r]) + 15) // 16 * 16
        src, dst = int(t["rsrc"][r]), int(t["rlds"][r])
        img[dst: dst + n16] = page[src: src + n16]
    return img


def _rows_at(img, at, bw, W):
    """bw staged rows of L = 1024/W words starting at LDS byte `at`, plus the row after them (pad or next)."""
    L = 1024 // W
    dt = W2DT[W]
    n = (bw + 1) * 128
    return np.frombuffer(img[at: at + n].tobytes(), dtype=dt).reshape(bw + 1, L).astype(np.uint64)


def emulate_tile(enc, t, out, dictionary=None):
    """Decodes tile t of the encoded buffer `enc` into the byte buffer `out` as the kernel does: the metadata comes
    from the page (one vector per lane), every staged byte from the tile's LDS image at the offsets the kernel
    computes (vector k's rows at ep_k - ep_0 in range 0, lane bases at range 1/2 + 128 k, run values at range 1 +
    r0_k - r0_0). Returns the error bits the kernel would set (1: RLE position clamped, 2: dictionary code
    clamped)."""
    fam, W = kind_parts(int(t["kind"]))
    pg = _Page(enc, t)
    dt = W2DT[W]
    img = stage_image(enc, t)
    rl = [int(x) for x in t["rlds"]]
    nv = int(t["nv"])
    v0 = int(t["v0"])
    meta = []
    for k in range(nv):  # the kernel's per-lane metadata loads
        v = v0 + k
        if fam == "unc":
            meta.append((W, 0, v * 128 * W, 0, 0))
        elif fam in ("ffor", "delta"):
            O, P = (1, 3) if fam == "delta" else (0, 2)
            meta.append((min(int(pg.b[pg.off[O] + v]), W), pg.u(pg.off[O + 1] + v * (W // 8), W // 8), pg.ep(v, P), 0, 0))
        elif fam == "rle":
            r1 = pg.ep(v + 1, 0) if v + 1 < int(t["vec_c"]) else pg.off[1] - pg.off[0]
            meta.append((min(int(pg.b[pg.off[2] + v]), 16), pg.u(pg.off[3] + 2 * v, 2), pg.ep(v, 4), pg.ep(v, 0), r1))
        elif fam == "dict1":
            meta.append((32, 0, v * 4096, 0, 0))
        else:
            O, P = (1, 3) if fam == "dict7" else (0, 2)
            meta.append((min(int(pg.b[pg.off[O] + v]), 32), pg.u(pg.off[O + 1] + 4 * v, 4), pg.ep(v, P), 0, 0))
    ep0, r00 = meta[0][2], meta[0][3]
    err = 0
    for k in range(nv):
        v = v0 + k
        bw, base, ep, r0, r1 = meta[k]
        at = rl[0] + ep - ep0
        if fam in ("unc", "ffor", "delta"):
            lb = None
            if fam == "delta":
                lb = np.frombuffer(img[rl[1] + 128 * k: rl[1] + 128 * k + 128].tobytes(), dtype=dt)
            vals = _int_vector(_rows_at(img, at, bw, W), bw, base, lb, W, fam == "delta").astype(dt)
        elif fam == "rle":
            n_runs = (r1 - r0) // (W // 8)
            runs = np.frombuffer(img[rl[1] + r0 - r00: rl[1] + r1 - r00].tobytes(), dtype=dt)
            lb = np.frombuffer(img[rl[2] + 128 * k: rl[2] + 128 * k + 128].tobytes(), dtype=np.uint16)
            idx = _int_vector(_rows_at(img, at, bw, 16), bw, base, lb, 16, True).astype(np.int64)
            if n_runs == 0 or (idx > n_runs - 1).any():
                err |= 1
            vals = runs[np.minimum(idx, max(n_runs - 1, 0))]
        else:
            var = int(fam[-1])
            lb = None
            if var == 7:
                lb = np.frombuffer(img[rl[1] + 128 * k: rl[1] + 128 * k + 128].tobytes(), dtype=np.uint32)
            codes = _int_vector(_rows_at(img, at, bw, 32), bw, base, lb, 32, var == 7).astype(np.int64)
            last = int(t["dict_n"]) - 1
            if (codes > last).any():
                err |= 2
            vals = np.asarray(dictionary)[np.minimum(codes, last)].astype(dt)
        o = int(t["out"]) + v * 128 * W
        out[o: o + 128 * W] = vals.view(np.uint8)
    return err


def fa_plan(columns, offsets, waves=0):
    """Encoded columns (bytes, dtype) at the given offsets -> (host buffer, tiles, LDS bytes per wave, wave runs)."""
    tiles, wave_lds, runs = fa.plan_work_list([c for c, _ in columns], [d for _, d in columns], offsets, waves)
    size = max(o + c.size for (c, _), o in zip(columns, offsets)) + 16
    host = np.zeros(size, dtype=np.uint8)
    for (c, _), o in zip(columns, offsets):
        host[o: o + c.size] = c
    return host, tiles, wave_lds, runs
