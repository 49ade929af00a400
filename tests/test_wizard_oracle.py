"""CPU checks of the wizard oracle: closed-form candidate sizes against the encoders' byte counts, sample statistics
against brute-force loops, and the MLP's analytic gradients against finite differences."""
import numpy as np
import pytest

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl
from oracle import wizard_oracle as wo


def edge_rowgroups():
    g = np.random.default_rng(5)
    yield "constant", np.full(wo.RG, 7, dtype=np.int64)
    yield "minus_one", np.full(wo.RG, -1, dtype=np.int64)  # all bits set
    x = g.integers(-3, 3, wo.RG).astype(np.int64)
    yield "small_signed", x
    yield "alternating", np.tile(np.array([0, np.iinfo(np.int64).max], dtype=np.int64), wo.RG // 2)
    yield "extremes", g.choice(np.array([np.iinfo(np.int64).min, -1, 0, np.iinfo(np.int64).max]), wo.RG)
    yield "unique_desc", np.arange(wo.RG, 0, -1, dtype=np.int64) * 1_000_003
    y = np.repeat(g.integers(0, 1 << 50, 5000), 14)[: wo.RG].astype(np.int64)
    yield "runs_many_distinct", np.concatenate([y, np.zeros(wo.RG - y.size, dtype=np.int64)])


@pytest.mark.parametrize("family", wl.WIZARD_FAMILIES)
def test_candidate_sizes_equal_encoded_sizes(oracle, family):
    for seed in range(2):
        x = wl.wizard_column(family, wo.RG, seed=seed + 10)
        s = wo.candidate_sizes(x)
        for k, eid in enumerate(wo.CANDIDATES):
            assert int(s[k]) == fa.encode_from_memory(x, eid).size == oracle.encode_column(x, eid).size, (family, eid)


def test_candidate_sizes_edge_cases(oracle):
    for name, x in edge_rowgroups():
        s = wo.candidate_sizes(x)
        for k, eid in enumerate(wo.CANDIDATES):
            assert int(s[k]) == oracle.encode_column(x, eid).size, (name, eid)


def test_hzl_size_full_width_chunk_ends_ffor_part(oracle):
    """A 1024-entry dictionary chunk spanning all 64 bits switches the rest of the dictionary to raw values."""
    g = np.random.default_rng(1)
    keys = np.concatenate([g.integers(0, 100, 1024), [np.iinfo(np.int64).min, np.iinfo(np.int64).max],
                           g.integers(0, 1 << 62, 3000)]).astype(np.int64)
    keys = np.unique(keys)[: 2048 + 100]
    g.shuffle(keys)
    x = np.concatenate([keys, np.repeat(keys[:1], wo.RG - keys.size)])
    s = wo.candidate_sizes(x)
    for k in (5, 6, 7):
        assert int(s[k]) == oracle.encode_column(x, wo.CANDIDATES[k]).size


def _brute_features(x):
    """Loop restatement of the integer statistics behind the features."""
    s = [list(map(int, x[v * 1024:(v + 1) * 1024].view(np.uint64))) for v in wo.SAMPLE_VECTORS]
    M = (1 << 64) - 1
    out = {"bwf": [], "bwd": [], "runs": [], "bwp": [], "bwc": [], "bwcd": []}
    seen, codes = {}, []
    for vec in s:
        for v in vec:
            if v not in seen:
                seen[v] = len(seen)
            codes.append(seen[v])
    sgn = lambda u: u - (1 << 64) if u >> 63 else u  # noqa: E731  (int64 columns: ANALYZE compares signed values)
    for i, vec in enumerate(s):
        sv = [sgn(v) for v in vec]
        out["bwf"].append((max(sv) - min(sv)).bit_length())
        d = [0 if o % 64 == 0 else sgn((vec[o] - vec[o - 1]) & M) for o in range(1024)]
        out["bwd"].append((max(d) - min(d)).bit_length())
        out["runs"].append(1 + sum(vec[o] != vec[o - 1] for o in range(1, 1024)))
        out["bwp"].append(int(any(vec[o] != vec[o - 1] for o in range(1, 1024) if o % 16)))
        c = codes[i * 1024:(i + 1) * 1024]
        out["bwc"].append((max(c) - min(c)).bit_length())
        cd = [0 if o % 32 == 0 else (c[o] - c[o - 1]) & 0xFFFFFFFF for o in range(1024)]
        out["bwcd"].append((max(cd) - min(cd)).bit_length())
    flat = [v for vec in s for v in vec]
    out["D"] = len(seen)
    out["top"] = max(flat.count(v) for v in set(flat)) if len(seen) < 64 else max(
        np.unique(np.array(flat, dtype=np.uint64), return_counts=True)[1])
    return out


@pytest.mark.parametrize("family", wl.WIZARD_FAMILIES)
def test_features_against_loops(family):
    x = wl.wizard_column(family, wo.RG, seed=3)
    F = wo.features(x)
    b = _brute_features(x)
    assert F.dtype == np.float32 and F.shape == (64,)
    G = 2 * len(wo.CANDIDATES)
    for i in range(4):
        assert F[G + 23 + i] * 64 == b["bwf"][i]
    assert F[G + 6] * 1024 == np.float32(sum(b["runs"])) / np.float32(4)
    assert F[G + 21] * 32 == np.float32(sum(b["bwc"])) / np.float32(4)
    assert F[G + 3] * 64 == np.float32(sum(b["bwd"])) / np.float32(4)
    assert F[G + 10] == np.float32(sum(b["bwp"])) / np.float32(4)
    assert F[G + 22] * 32 == np.float32(sum(b["bwcd"])) / np.float32(4)
    assert F[G + 11] * 4096 == b["D"]
    assert F[G + 14] * 4096 == b["top"]
    assert F[63] == 1.0
    # the sample-size estimates of the non-dictionary candidates use the closed forms of candidate_sizes
    assert abs(F[1] * wo.RG * 8 - (8 + 12 * 64 + 64 * (9 + 128 * np.mean(b["bwf"])))) < 1e-2 * wo.RG


def test_mlp_gradients_match_finite_differences():
    g = np.random.default_rng(0)
    p = (g.standard_normal(wo.N_PARAMS) * 0.1).astype(np.float64)
    X = g.standard_normal((6, 64))
    y = g.integers(0, 8, 6)
    loss, grad, _ = wo.mlp_forward_backward(p, X, y, 8, rounding=False)
    idx = np.concatenate([g.choice(wo.N_PARAMS, 60, replace=False),
                          [wo.N_PARAMS - 32, wo.N_PARAMS - 25, 256 * 64 + 3]])  # includes b3 and b1 entries
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (wo.mlp_loss(p + e, X, y, 8) - wo.mlp_loss(p - e, X, y, 8)) / 2e-6
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])
    # padded classes get no gradient
    W3t_g = grad[256 * 64 + 256 + 256 * 256 + 256:][: 32 * 256].reshape(32, 256)
    assert np.all(W3t_g[8:] == 0)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5e-3, 3.0e38], dtype=np.float32)
    r = wo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625  # ties to even, above-half rounds up
    assert abs(r[3] - -2.5e-3) < 2e-5


# -- structural wizard (Wizard::Spell) ----------------------------------------------------------------------------

def oracle_columns(columns):
    """Columns as the oracle takes them: (type name, Python values)."""
    out = []
    for c in columns:
        if isinstance(c, tuple):
            offs, data = c
            b = data.tobytes()
            out.append(("str", [b[int(offs[k]): int(offs[k + 1])] for k in range(len(offs) - 1)]))
        else:
            out.append((c.dtype.name, c.tolist()))
    return out


def test_spell_example_1_plan():
    cols, kinds, refs = wl.spell_example_1()
    assert wo.spell(oracle_columns(cols)) == (kinds, refs)


@pytest.mark.parametrize("case", list(wl.spell_edge_tables()), ids=lambda c: c[0])
def test_spell_edge_tables(case):
    _, cols, kinds, refs = case
    assert wo.spell(oracle_columns(cols)) == (kinds, refs)


def test_spell_operator_lists():
    from fastlanes_amd import spell

    _, cols, kinds, refs = next(c for c in wl.spell_edge_tables() if c[0] == "several_seconds")
    ops = spell.operators(kinds, refs)
    assert ops[0] == ["DICT", "SUB_EXPRESSION"] * 3 and ops[4] == ["DICT", "SUB_EXPRESSION"]
    assert ops[1] == ["FIXME"]
    cols, kinds, refs = wl.spell_example_1()
    ops = spell.operators(kinds, refs)
    assert ops[:3] == [["CONSTANT"], ["CONSTANT"], ["ZERO_COPY"]] and ops[5] == ["EQUAL"]


def _patch_plan_loops(v):
    """plan_patch as literal loops over the sorted vector (the host encoder's search): (bw, base)."""
    sv = sorted(int(a) for a in v.view(np.uint64))
    M = (1 << 64) - 1
    best_size, best_i, best_bw, found = 128 * 64, 0, 64, False
    for bw in range(64):
        rng_ = (1 << bw) - 1
        j = 0
        for i in range(1024):
            if i and sv[i] == sv[i - 1]:
                continue
            if bw == 0:
                inw = 0
                if sv[i] == 0:
                    while inw < 1024 and sv[inw] == 0:
                        inw += 1
            else:
                upper = M if sv[i] > M - rng_ else sv[i] + rng_
                j = max(j, i)
                while j < 1024 and sv[j] <= upper:
                    j += 1
                inw = j - i
            exc = 1024 - inw
            size = 128 * bw + 8 * exc
            if exc > wo.EXC_LIMIT:
                continue
            if size < best_size or (size == best_size and found and (i < best_i or (i == best_i and bw > best_bw))):
                best_size, best_i, best_bw, found = size, i, bw, True
    return (best_bw, sv[best_i]) if found else (64, 0)


@pytest.mark.parametrize("case", ["outliers", "uniform", "zeros", "wide", "ties"])
def test_patch_plan_against_loops(case):
    """The oracle's vectorized patch plan equals the literal search on vectors built to exercise its branches: a
    narrow body with outliers, uniform values, mostly zeros (bw = 0 windows), full 64-bit values, equal-size ties."""
    g = np.random.default_rng(hash(case) % 1000)
    if case == "outliers":
        v = g.integers(1000, 1100, 1024).astype(np.uint64)
        v[g.choice(1024, 37, replace=False)] = g.integers(0, 1 << 62, 37).astype(np.uint64)
    elif case == "uniform":
        v = g.integers(0, 1 << 20, 1024).astype(np.uint64)
    elif case == "zeros":
        v = np.zeros(1024, np.uint64)
        v[g.choice(1024, 60, replace=False)] = g.integers(1, 1 << 40, 60).astype(np.uint64)
    elif case == "wide":
        v = g.integers(0, (1 << 64) - 1, 1024, dtype=np.uint64, endpoint=True)
    else:
        v = np.repeat(np.arange(16, dtype=np.uint64) * 7, 64)
        v[:3] = np.uint64((1 << 64) - 1)
    assert wo.patch_plan(v) == _patch_plan_loops(v)

