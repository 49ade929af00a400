"""Seeded synthetic column families.

These are the inputs of the decode benchmarks and of the encoding wizard's training set. All generators
use numpy's PCG64 bit generator with an explicit seed, so every host (including the GPU host) produces the
same bytes for the same arguments.
"""
from __future__ import annotations

import numpy as np

W2U = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def config1_values(n: int = 1 << 20, seed: int = 42) -> np.ndarray:
    """FFOR u32 workload: 1,000,000 + U[0, 2^7), bit width 7."""
    return (np.uint32(1_000_000) + rng(seed).integers(0, 128, n, dtype=np.uint32)).astype(np.uint32)


def sweep_case(W: int, w: int, n_vec: int, seed: int = 0):
    """Bit-width sweep case: values U[0, 2^w) and per-vector FFOR bases U[0, 2^(W-w)] (inclusive; at w = 0 the
    bound 2^W is clamped to 2^W - 1). Returns (packed-domain values, bases); FFOR values are value + base."""
    g = rng(seed * 1000 + W * 100 + w)
    dt = W2U[W]
    hi = (1 << w) if w < 64 else None
    if w == 0:
        vals = np.zeros(n_vec * 1024, dtype=dt)
    elif w == W:
        vals = g.integers(0, np.iinfo(dt).max, n_vec * 1024, dtype=dt, endpoint=True)
    else:
        vals = g.integers(0, hi, n_vec * 1024, dtype=np.uint64).astype(dt)
    top = (1 << (W - w)) if W - w < 64 else (1 << 64) - 1
    top = min(top, (1 << W) - 1)
    bases = g.integers(0, top, n_vec, dtype=np.uint64, endpoint=True).astype(dt)
    return vals, bases


def timestamps(n: int, seed: int = 3) -> np.ndarray:
    """Sorted int64 timestamps starting at 1.7e18 with gaps U[1, 100]."""
    gaps = rng(seed).integers(1, 101, n, dtype=np.int64)
    gaps[0] = 0
    return np.int64(1_700_000_000_000_000_000) + np.cumsum(gaps, dtype=np.int64)


def geometric_runs(n: int, mean: float = 16.0, seed: int = 4, value_bits: int = 20,
                   dtype=np.uint32) -> np.ndarray:
    """Runs with Geometric(1/mean) lengths and values U[0, 2^value_bits)."""
    g = rng(seed)
    lens = g.geometric(1.0 / mean, size=n // max(int(mean), 1) * 2 + 16)
    while lens.sum() < n:
        lens = np.concatenate([lens, g.geometric(1.0 / mean, size=lens.size)])
    vals = g.integers(0, 1 << value_bits, lens.size, dtype=np.uint64).astype(dtype)
    return np.repeat(vals, lens)[:n].copy()


def zipf_codes(n: int, n_keys: int = 4096, s: float = 1.1, seed: int = 5, dtype=np.int32) -> np.ndarray:
    """Categorical column: n_keys distinct random keys drawn with Zipf(s) over key rank."""
    g = rng(seed)
    info = np.iinfo(dtype)
    keys = g.choice(np.arange(info.min // 2, info.max // 2, dtype=np.int64), size=n_keys, replace=False)
    p = 1.0 / np.arange(1, n_keys + 1, dtype=np.float64) ** s
    p /= p.sum()
    ranks = g.choice(n_keys, size=n, p=p)
    return keys[ranks].astype(dtype)


def vocabulary(n_words: int = 2000, seed: int = 11) -> list[bytes]:
    g = rng(seed)
    letters = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz", dtype=np.uint8)
    words, seen = [], set()
    while len(words) < n_words:
        ln = int(g.integers(3, 11))
        w = letters[g.integers(0, 26, ln)].tobytes()
        if w not in seen:
            seen.add(w)
            words.append(w)
    return words


def vocab_strings(n: int, seed: int = 7, min_len: int = 8, max_len: int = 64, n_words: int = 2000,
                  words_per_string: tuple[int, int] | None = None):
    """Strings of a 2000-word lowercase vocabulary. With words_per_string=None each string has a length
    U[min_len, max_len] (words joined by spaces, cut to length); otherwise it has U[lo, hi] words.
    Returns (offsets uint32[n+1], bytes uint8)."""
    g = rng(seed)
    vocab = vocabulary(n_words)
    idx = g.integers(0, len(vocab), size=n * 12)
    out, offs, k = bytearray(), np.zeros(n + 1, dtype=np.uint32), 0
    for i in range(n):
        if words_per_string is None:
            target = int(g.integers(min_len, max_len + 1))
            s = bytearray()
            while len(s) < target:
                if s:
                    s += b" "
                s += vocab[idx[k % idx.size]]
                k += 1
            s = s[:target]
        else:
            cnt = int(g.integers(words_per_string[0], words_per_string[1] + 1))
            s = b" ".join(vocab[idx[(k + j) % idx.size]] for j in range(cnt))
            k += cnt
        out += s
        offs[i + 1] = len(out)
    return offs, np.frombuffer(bytes(out), dtype=np.uint8).copy()


def for_with_outliers(n: int, width: int = 8, outlier_frac: float = 0.01, seed: int = 6) -> np.ndarray:
    g = rng(seed)
    x = np.int64(123_456) + g.integers(0, 1 << width, n, dtype=np.int64)
    m = g.random(n) < outlier_frac
    x[m] = g.integers(0, 1 << 40, int(m.sum()), dtype=np.int64)
    return x


def patch_column(dtype, n: int, seed: int = 8) -> np.ndarray:
    """Data for the patched FFOR variants (32/64-bit): per 1024-value vector a body of 0-20 bits above a random base
    with up to 120 outliers anywhere in the type's range, and every eighth vector a corner case in turn: constant,
    all zero but three values far apart (a bit width of 0 makes LL_PATCH link every position in between), full-range
    values (no patch pays), exactly 100 and exactly 101 outliers."""
    dt = np.dtype(dtype)
    bits = dt.itemsize * 8
    udt = np.uint64 if bits == 64 else np.uint32
    umax = (1 << bits) - 1
    g = rng(seed)
    nv = n // 1024
    out = np.empty((nv, 1024), dtype=udt)
    for v in range(nv):
        row = out[v]
        corner = (v // 8) % 5 if v % 8 == 7 else -1
        if corner == 0:
            row[:] = g.integers(0, umax, dtype=udt, endpoint=True)
        elif corner == 1:
            row[:] = 0
            row[[3, 500, 1020]] = g.integers(1, umax, 3, dtype=udt, endpoint=True)
        elif corner == 2:
            row[:] = g.integers(0, umax, 1024, dtype=udt, endpoint=True)
        else:
            bw = int(g.integers(0, 21))
            base = int(g.integers(0, umax - (1 << bw) + 1, dtype=udt, endpoint=True))
            row[:] = (base + g.integers(0, 1 << bw, 1024, dtype=np.uint64)).astype(udt)
            k = 100 if corner == 3 else 101 if corner == 4 else int(g.integers(0, 121))
            pos = g.choice(1024, k, replace=False)
            row[pos] = g.integers(0, umax, k, dtype=udt, endpoint=True)
    return out.reshape(-1).view(dt)


def vocab_strings_fast(n: int, seed: int = 7, min_len: int = 8, max_len: int = 64, n_words: int = 2000):
    """Vectorized variant for large n: a text of space-separated vocabulary words is cut into consecutive strings
    of length U[min_len, max_len]. Returns (offsets uint32[n+1], bytes uint8)."""
    g = rng(seed)
    vocab = vocabulary(n_words)
    lens = g.integers(min_len, max_len + 1, n).astype(np.int64)
    total = int(lens.sum())
    wl = np.array([len(w) + 1 for w in vocab], dtype=np.int64)  # word + space
    wbytes = np.frombuffer(b"".join(w + b" " for w in vocab), dtype=np.uint8)
    wstart = np.concatenate([[0], np.cumsum(wl)[:-1]])
    n_words_needed = int(total / wl.mean() * 1.1) + 16
    idx = g.integers(0, len(vocab), n_words_needed)
    pieces_len = wl[idx]
    while pieces_len.sum() < total:
        more = g.integers(0, len(vocab), n_words_needed // 10 + 16)
        idx = np.concatenate([idx, more])
        pieces_len = wl[idx]
    # the text is the concatenation of the chosen word pieces (+ space), cut to `total` bytes
    ends = np.cumsum(pieces_len)
    cut = int(np.searchsorted(ends, total)) + 1
    idx, pieces_len, ends = idx[:cut], pieces_len[:cut], ends[:cut]
    rel = np.repeat(wstart[idx] - (ends - pieces_len), pieces_len)
    text = wbytes[rel[:total] + np.arange(total, dtype=np.int64)]
    offs = np.zeros(n + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    if offs[-1] >= 1 << 32:
        raise ValueError("string column too large for 32-bit offsets")
    return offs.astype(np.uint32), text


# expression ids of the mixed row-group (include/fls/expression/expression_pool, i64 pool numbering)
ROWGROUP_EXPRS = {"bitpack": 1, "ffor": 1, "delta": 11, "rle": 6, "dict": 17}


def rowgroup_column(kind: str, W: int, n: int, seed: int, w: int | None = None) -> np.ndarray:
    """One column of the mixed row-group: `kind` in bitpack / ffor / delta / rle / dict, element width W bits."""
    g = rng(seed)
    dt = W2U[W]
    if kind == "bitpack":  # uniform in [0, 2^w), w in 1..min(32, W)
        w = w if w is not None else int(g.integers(1, min(32, W) + 1))
        return g.integers(0, 1 << w, n, dtype=np.uint64).astype(dt)
    if kind == "ffor":  # a large base plus uniform in [0, 2^w)
        w = w if w is not None else int(g.integers(1, min(32, W - 1) + 1))
        base = int(g.integers(0, (1 << W) - (1 << w), dtype=np.uint64, endpoint=True))
        return (np.uint64(base) + g.integers(0, 1 << w, n, dtype=np.uint64)).astype(dt)
    if kind == "delta":  # sorted (wrapping in narrow types) with gaps uniform in [0, 2^gw)
        gw = int(g.integers(1, min(12, W - 1) + 1))
        x = np.cumsum(g.integers(0, 1 << gw, n, dtype=np.uint64), dtype=np.uint64) + np.uint64(g.integers(0, 1 << 30))
        return x.astype(dt)
    if kind == "rle":  # geometric runs of mean 16
        return geometric_runs(n, 16.0, seed=seed, value_bits=min(W, 20), dtype=dt)
    if kind == "dict":  # few distinct wide values
        keys = g.integers(0, np.iinfo(dt).max, int(g.integers(16, 4097)), dtype=dt, endpoint=True)
        return keys[g.integers(0, keys.size, n)]
    raise ValueError(kind)


def rowgroup_spec(n_columns: int = 64, seed: int = 17, min_log2: int = 20, max_log2: int = 26):
    """Mixed row-group: per column (W, kind, n) with W in {8, 16, 32, 64}, kind in bitpack / ffor / delta / rle /
    dict and n uniform in [2^min_log2, 2^max_log2] rounded to a multiple of 1024."""
    g = rng(seed)
    kinds = list(ROWGROUP_EXPRS)
    out = []
    for c in range(n_columns):
        W = int(g.choice([8, 16, 32, 64]))
        kind = kinds[int(g.integers(0, len(kinds)))]
        n = int(g.integers(1 << min_log2, (1 << max_log2) + 1))
        n = max(1024, n // 1024 * 1024)
        out.append((W, kind, n, seed * 1000 + c))
    return out


WIZARD_FAMILIES = ("uniform", "sorted", "runs", "zipf", "outliers")


def wizard_column(family: str, n: int = 1 << 20, seed: int = 0) -> np.ndarray:
    """int64 column of one of the encoding-wizard families: uniform ints of width 1..40; sorted sequences; geometric
    runs; Zipf categoricals with 16..65536 keys; FOR-able data with 1% uniform outliers."""
    g = rng(seed)
    if family == "uniform":
        w = int(g.integers(1, 41))
        return g.integers(0, 1 << w, n, dtype=np.int64)
    if family == "sorted":
        gap = int(1 << int(g.integers(0, 17)))
        start = int(g.integers(0, 1 << 40))
        return np.int64(start) + np.cumsum(g.integers(0, gap + 1, n, dtype=np.int64))
    if family == "runs":
        mean = float(2 ** g.uniform(1, 7))
        return geometric_runs(n, mean, seed=seed, value_bits=int(g.integers(4, 33)), dtype=np.int64)
    if family == "zipf":
        n_keys = int(2 ** g.uniform(4, 16))
        kw = int(g.integers(16, 48))
        keys = g.choice(1 << kw, size=n_keys, replace=False).astype(np.int64)
        s = float(g.uniform(0.8, 1.6))
        p = 1.0 / np.arange(1, n_keys + 1, dtype=np.float64) ** s
        p /= p.sum()
        return keys[g.choice(n_keys, size=n, p=p)]
    if family == "outliers":
        w = int(g.integers(1, 25))
        x = np.int64(g.integers(0, 1 << 40)) + g.integers(0, 1 << w, n, dtype=np.int64)
        m = g.random(n) < 0.01
        x[m] = g.integers(0, 1 << 40, int(m.sum()), dtype=np.int64)
        return x
    raise ValueError(family)


def wizard_columns(n_columns: int = 512, n: int = 1 << 20, seed: int = 99):
    """The wizard data set: column c belongs to family c mod 5, with its own seed."""
    return [(WIZARD_FAMILIES[c % len(WIZARD_FAMILIES)], wizard_column(WIZARD_FAMILIES[c % len(WIZARD_FAMILIES)], n,
                                                                       seed * 100003 + c)) for c in range(n_columns)]


# -- structural wizard tables -----------------------------------------------------------------------------------

def _strs(items) -> tuple[np.ndarray, np.ndarray]:
    """(offsets u64, bytes u8) of a list of bytes objects."""
    lens = np.fromiter((len(b) for b in items), dtype=np.uint64, count=len(items))
    offs = np.zeros(len(items) + 1, dtype=np.uint64)
    np.cumsum(lens, out=offs[1:])
    return offs, np.frombuffer(b"".join(items), dtype=np.uint8).copy()


SPELL_EXAMPLE_1 = ["CONSTANT_INT", "CONSTANT_STR", "EQUAL_GROUP_0_COL_0", "RANGE_0_1000", "NUMBER_AS_STRING",
                   "EQUAL_GROUP_0_COL_1", "EQUAL_GROUP_0_COL_2", "MAP1T1_GROUP_0_COL_0", "MAP1T1_GROUP_0_COL_1"]


def spell_example_1(n: int = 8192, seed: int = 1):
    """A table with the columns of the reference's generated example (data/generated/example_1/schema.json), filled
    as their names say: a constant integer and string, a group of three equal integer columns, integers in
    [0, 1000), numbers written as strings, and a string key column with an integer column that is a function of it.
    Returns (columns, expected kinds, expected refs) -- the plan Wizard::Spell makes for such data."""
    g = rng(seed)
    eq = g.integers(-(1 << 62), 1 << 62, n, dtype=np.int64)
    keys = g.integers(0, 100, n)
    cols = [
        np.full(n, 7, dtype=np.int64),
        _strs([b"fastlanes"] * n),
        eq,
        g.integers(0, 1000, n).astype(np.int64),
        _strs([str(v).encode() for v in g.integers(0, 1 << 40, n)]),
        eq.copy(),
        eq.copy(),
        _strs([b"key_%d" % k for k in keys]),
        (1000 + 7 * keys).astype(np.int64),
    ]
    kinds = ["CONSTANT", "CONSTANT", "EXPRESSION", "EXPRESSION", "EXPRESSION", "EQUAL", "EQUAL", "MAP_SOURCE",
             "MAP_TARGET"]
    refs = [-1, -1, -1, -1, -1, 2, 2, 8, 7]
    return cols, kinds, refs


def spell_edge_tables():
    """Tables whose plan follows from their construction: the 25% distinct boundary, NaN and signed zeros, a type
    mismatch, one first with several seconds, a second already taken. Yields (name, columns, kinds, refs)."""
    n = 4096
    i = np.arange(n)
    g = rng(5)
    c4 = g.standard_normal(n)
    c4[::7] = 0.0
    c4[3::7] = -0.0
    c5 = c4.copy()
    c5[::7] = -0.0
    c5[3::7] = 0.0
    c6 = ((i * 7919) % 5).astype(np.float64)
    c6[5] = np.nan
    yield ("boundary_nan_zero",
           [(i % 1024).astype(np.int32), ((i % 1024) * 3 + 1).astype(np.int64), (i % 1025).astype(np.uint16),
            (i % 1025).astype(np.float64), c4, c5, c6, c6.copy()],
           ["MAP_SOURCE", "MAP_TARGET", "EXPRESSION", "EXPRESSION", "EXPRESSION", "EQUAL", "MAP_SOURCE", "MAP_TARGET"],
           [1, 0, -1, -1, -1, 4, 7, 6])
    n = 2048
    i = np.arange(n)
    yield ("several_seconds",
           [(i % 16).astype(np.int64), (i % 16).astype(np.uint64), _strs([b"s%d" % v for v in i % 16]),
            (i % 8).astype(np.int8), (i % 32).astype(np.int64), ((i % 32) // 2).astype(np.int64)],
           ["MAP_SOURCE", "MAP_TARGET", "MAP_TARGET", "MAP_TARGET", "MAP_SOURCE", "MAP_TARGET"],
           [1, 0, 0, 0, 5, 4])
    yield ("constants", [np.full(n, 5, np.int64), np.full(n, 5, np.int64), _strs([b""] * n)],
           ["CONSTANT", "CONSTANT", "CONSTANT"], [-1, -1, -1])
    yield ("one_row", [np.array([3], np.int64), np.array([4.0]), _strs([b"x"])],
           ["CONSTANT", "CONSTANT", "CONSTANT"], [-1, -1, -1])
    yield ("empty", [np.zeros(0, np.int64), np.zeros(0, np.uint8)], ["CONSTANT", "CONSTANT"], [-1, -1])


def delta_patch_column(dtype, n: int, seed: int = 9) -> np.ndarray:
    """Data for the patched DELTA variants (32/64-bit): per vector a running sum of small steps (0 to 2^k - 1 for a
    random k <= 12) with up to 120 jumps anywhere in the type's range, so the lane differences are a narrow body with
    outliers; every eighth vector a corner case in turn: constant, strictly increasing by one, a running sum with
    exactly 100 and exactly 101 jumps, and differences over the full range."""
    dt = np.dtype(dtype)
    bits = dt.itemsize * 8
    udt = np.uint64 if bits == 64 else np.uint32
    g = rng(seed)
    nv = n // 1024
    out = np.empty((nv, 1024), dtype=np.uint64)
    mask = np.uint64((1 << bits) - 1)
    for v in range(nv):
        corner = (v // 8) % 5 if v % 8 == 7 else -1
        start = np.uint64(int(g.integers(0, 1 << (bits - 1))))
        if corner == 0:
            steps = np.zeros(1024, dtype=np.uint64)
        elif corner == 1:
            steps = np.ones(1024, dtype=np.uint64)
        elif corner == 4:
            steps = g.integers(0, 1 << 63, 1024, dtype=np.uint64) * np.uint64(2) + g.integers(0, 2, 1024, dtype=np.uint64)
        else:
            k = int(g.integers(0, 13))
            steps = g.integers(0, 1 << k, 1024, dtype=np.uint64)
            nj = 100 if corner == 2 else 101 if corner == 3 else int(g.integers(0, 121))
            pos = g.choice(1024, nj, replace=False)
            steps[pos] = g.integers(0, 1 << 62, nj, dtype=np.uint64) * np.uint64(3)
        with np.errstate(over="ignore"):
            out[v] = (start + np.cumsum(steps, dtype=np.uint64)) & mask
    return out.astype(udt).reshape(-1).view(dt)


def dict_rare_keys(n: int, dtype=np.int64, n_common: int = 200, n_rare: int = 40, rare_frac: float = 0.01,
                   seed: int = 12) -> np.ndarray:
    """Dictionary data whose codes suit the patched dictionary variants: n_common keys drawn uniformly (codes below
    n_common, in first-occurrence or frequency order alike) and, at a rate of rare_frac, n_rare other keys that first
    appear late in the column (high codes: the exceptions of a narrow code width); runs of one key give the DELTA
    code variants narrow differences."""
    info = np.iinfo(dtype)
    g = rng(seed)
    keys = g.choice(np.arange(info.min, info.max, max(1, (int(info.max) - int(info.min)) // (4 * (n_common + n_rare))),
                              dtype=np.int64)[: 4 * (n_common + n_rare)], n_common + n_rare, replace=False)
    idx = geometric_runs(n, 6.0, seed=seed + 1, value_bits=16, dtype=np.int64) % n_common
    rare = g.random(n) < rare_frac
    rare[: n // 2] = False
    idx[rare] = n_common + g.integers(0, n_rare, int(rare.sum()))
    return keys[idx].astype(dtype)


def dict_runs(n: int, dtype=np.int64, n_keys: int = 300, mean: float = 12.0, seed: int = 6) -> np.ndarray:
    """Runs (Geometric(1/mean) lengths) over n_keys random keys of the type: the shape DICT_RLE compresses."""
    info = np.iinfo(dtype)
    keys = rng(seed).integers(info.min, info.max, n_keys, dtype=dtype, endpoint=True)
    idx = geometric_runs(n, mean, seed=seed + 1, value_bits=16, dtype=np.int64) % n_keys
    return keys[idx]
