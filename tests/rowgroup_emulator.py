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


def claim_loop(n_tiles, n_waves, rng):
    """The kernel's dispatch: every wave takes ids from one counter until an id is past the list. Waves run in a
    random interleaving; returns the ids each wave decoded and the number of counter increments."""
    counter, done, claims = 0, [False] * n_waves, [[] for _ in range(n_waves)]
    steps = 0
    while not all(done):
        w = int(rng.integers(0, n_waves))
        if done[w]:
            continue
        steps += 1
        assert steps <= 10 * (n_tiles + n_waves), "claim loop did not end"
        tid, counter = counter, counter + 1
        if tid >= n_tiles:
            done[w] = True
        else:
            claims[w].append(tid)
    return claims, counter


class _Page:
    def __init__(self, enc, t):
        self.b = enc[int(t["page"]):]
        self.arr_c = int(t["arr_c"])
        self.off = [int(x) for x in t["off"]]

    def u(self, at, n):
        return int.from_bytes(self.b[at: at + n].tobytes(), "little")

    def ep(self, v, buf):
        return 0 if v == 0 else self.u(8 + 4 * self.arr_c + 4 * ((v - 1) * self.arr_c + buf), 4)


def _rows(pg, at, bw, W):
    """Staged rows of a vector: bw rows of L = 1024/W words, plus the zero pad row the kernel never keeps."""
    L = 1024 // W
    dt = W2DT[W]
    rows = np.zeros((bw + 1, L), dtype=np.uint64)
    if bw:
        rows[:bw] = np.frombuffer(pg.b[at: at + 128 * bw].tobytes(), dtype=dt).reshape(bw, L).astype(np.uint64)
    return rows


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


def emulate_tile(enc, t, out, dictionary=None):
    """Decodes tile t of the encoded buffer `enc` into the byte buffer `out` as the kernel would. Returns the error
    bits the kernel would set (1: RLE position clamped, 2: dictionary code clamped)."""
    fam, W = kind_parts(int(t["kind"]))
    pg = _Page(enc, t)
    dt = W2DT[W]
    err = 0
    for k in range(int(t["nv"])):
        v = int(t["v0"]) + k
        if fam in ("unc", "ffor", "delta"):
            O = 1 if fam == "delta" else 0
            P = 0 if fam == "unc" else (2 if fam == "ffor" else 3)
            if fam == "unc":
                bw, base, at = W, 0, pg.off[0] + v * 128 * W
            else:
                bw = int(pg.b[pg.off[O] + v])
                base = pg.u(pg.off[O + 1] + v * (W // 8), W // 8)
                at = pg.off[P] + pg.ep(v, P)
            lb = None
            if fam == "delta":
                lb = np.frombuffer(pg.b[pg.off[0] + 128 * v: pg.off[0] + 128 * v + 128].tobytes(), dtype=dt)
            vals = _int_vector(_rows(pg, at, bw, W), bw, base, lb, W, fam == "delta").astype(dt)
        elif fam == "rle":
            bw = int(pg.b[pg.off[2] + v])
            base = pg.u(pg.off[3] + 2 * v, 2)
            r0 = pg.ep(v, 0)
            r1 = pg.ep(v + 1, 0) if v + 1 < int(t["vec_c"]) else pg.off[1] - pg.off[0]
            runs = np.frombuffer(pg.b[pg.off[0] + r0: pg.off[0] + r1].tobytes(), dtype=dt)
            lb = np.frombuffer(pg.b[pg.off[1] + 128 * v: pg.off[1] + 128 * v + 128].tobytes(), dtype=np.uint16)
            idx = _int_vector(_rows(pg, pg.off[4] + pg.ep(v, 4), bw, 16), bw, base, lb, 16, True).astype(np.int64)
            if runs.size == 0 or (idx > runs.size - 1).any():
                err |= 1
            vals = runs[np.minimum(idx, max(runs.size - 1, 0))]
        else:
            var = int(fam[-1])
            if var == 1:
                bw, base, at, lb = 32, 0, pg.off[0] + 4096 * v, None
            else:
                O = 1 if var == 7 else 0
                P = 3 if var == 7 else 2
                bw = int(pg.b[pg.off[O] + v])
                base = pg.u(pg.off[O + 1] + 4 * v, 4)
                at = pg.off[P] + pg.ep(v, P)
                lb = np.frombuffer(pg.b[pg.off[0] + 128 * v: pg.off[0] + 128 * v + 128].tobytes(),
                                   dtype=np.uint32) if var == 7 else None
            codes = _int_vector(_rows(pg, at, bw, 32), bw, base, lb, 32, var == 7).astype(np.int64)
            last = int(t["dict_n"]) - 1
            if (codes > last).any():
                err |= 2
            vals = np.asarray(dictionary)[np.minimum(codes, last)].astype(dt)
        o = int(t["out"]) + v * 128 * W
        out[o: o + 128 * W] = vals.view(np.uint8)
    return err


def fa_plan(columns, offsets):
    """Encoded columns (bytes, dtype) at the given offsets -> (host buffer, tiles, LDS bytes per wave)."""
    tiles, wave_lds = fa.plan_work_list([c for c, _ in columns], [d for _, d in columns], offsets)
    size = max(o + c.size for (c, _), o in zip(columns, offsets)) + 16
    host = np.zeros(size, dtype=np.uint8)
    for (c, _), o in zip(columns, offsets):
        host[o: o + c.size] = c
    return host, tiles, wave_lds
