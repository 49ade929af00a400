"""Structural wizard on the GPU (fls_wizard_spell) against the oracle's restatement of Wizard::Spell and against
tables built to a known plan."""
import numpy as np
import pytest

from fastlanes_amd import workloads as wl
from oracle import wizard_oracle as wo
from tests.test_wizard_oracle import oracle_columns

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spell():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fastlanes_amd import spell as sp

    return sp


def _plan(sp, cols):
    res = sp.spell(cols)
    return [r["kind"] for r in res], [r["ref"] for r in res]


def test_example_1(spell):
    print("spell example_1", flush=True)
    cols, kinds, refs = wl.spell_example_1()
    assert _plan(spell, cols) == (kinds, refs)


@pytest.mark.parametrize("case", list(wl.spell_edge_tables()), ids=lambda c: c[0])
def test_edge_tables(spell, case):
    name, cols, kinds, refs = case
    print(f"spell {name}", flush=True)
    assert _plan(spell, cols) == (kinds, refs)


def _random_table(n, n_cols, seed):
    """Columns drawn from: constants, copies of earlier columns (same type, or another type), functions of an
    earlier low-cardinality column, independent low- and high-cardinality columns, strings and doubles."""
    g = np.random.default_rng(seed)
    cols = []
    for c in range(n_cols):
        kind = int(g.integers(0, 8)) if cols else 3
        if kind == 0:
            cols.append(np.full(n, g.integers(0, 100), dtype=np.int32))
        elif kind == 1:
            src = cols[int(g.integers(0, len(cols)))]
            cols.append(src if isinstance(src, tuple) else src.copy())
        elif kind == 2:
            src = cols[int(g.integers(0, len(cols)))]
            if isinstance(src, tuple):
                cols.append(np.zeros(n, np.int64))
            else:
                u, inv = np.unique(src, return_inverse=True)
                table = g.integers(0, 1 << 30, u.size)
                cols.append(table[inv].astype(np.int64))
        elif kind == 3:
            cols.append(g.integers(0, int(g.choice([4, 50, 2000])), n).astype(np.int16))
        elif kind == 4:
            cols.append(g.integers(0, 1 << 62, n, dtype=np.int64))
        elif kind == 5:
            v = g.integers(0, int(g.choice([10, 100000])), n)
            cols.append(wl._strs([b"w%d" % x for x in v]))
        elif kind == 6:
            x = g.integers(0, 20, n).astype(np.float64)
            if g.random() < 0.5:
                x[int(g.integers(0, n))] = np.nan
            cols.append(x)
        else:
            src = cols[int(g.integers(0, len(cols)))]
            cols.append(src.astype(np.uint64) if not isinstance(src, tuple) and src.dtype != np.float64 else src)
    return cols


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_tables_match_oracle(spell, seed):
    print(f"spell random table seed={seed}", flush=True)
    cols = _random_table(1 << 14, 20, seed)
    want = wo.spell(oracle_columns(cols))
    assert _plan(spell, cols) == want


def test_large_rowgroup(spell):
    """64 columns of 2^20 rows with a plan known by construction (no oracle run at this size)."""
    print("spell 64 x 2^20", flush=True)
    n = 1 << 20
    g = np.random.default_rng(4)
    i = np.arange(n)
    cols, kinds, refs = [], [], []
    for blk in range(16):
        hi = g.integers(0, 1 << 62, n, dtype=np.int64)
        key = ((i * (blk + 3)) % (1000 + blk)).astype(np.int32)
        cols += [hi, hi.copy(), key, (key.astype(np.int64) * 11 + blk)]
        b = 4 * blk
        kinds += ["EXPRESSION", "EQUAL", "MAP_SOURCE", "MAP_TARGET"]
        refs += [-1, b, b + 3, b + 2]
    assert _plan(spell, cols) == (kinds, refs)
