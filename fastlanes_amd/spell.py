"""Structural wizard: the plan Wizard::Spell (src/wizard/wizard.cpp:19-44) makes for a row-group, computed on the GPU.

Per column, in the reference's order: CONSTANT (all values equal), EQUAL (same type and values as an earlier
column), a one-to-one map (a first column with at most 25% distinct values whose every value always meets the same
value of a later column: DICT + SUB_EXPRESSION on the first, FIXME on the second), else ZERO_COPY (an expression of
its own). `spell` returns one entry per column with the kind, the referenced column and the operator list the
reference's footer would hold (`ColumnMetadata.encoding_rpn.operators`, include/fls/footer/rowgroup_footer.hpp).

Columns are numpy arrays (integers or float64) or (offsets, bytes) pairs for strings, all of one length.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import fastlanes_amd as fa

KINDS = ["EXPRESSION", "CONSTANT", "EQUAL", "MAP_SOURCE", "MAP_TARGET"]
_STR = 10  # data_t STR


class SpellColumn(C.Structure):
    _fields_ = [("type", C.c_uint8), ("values", C.c_void_p), ("offsets", C.c_void_p)]


def _lib():
    L = fa.lib()
    if not getattr(L, "_spell_bound", False):
        L.fls_wizard_spell.argtypes = [C.c_uint32, C.POINTER(SpellColumn), C.c_uint64, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.c_void_p]
        L._spell_bound = True
    return L


def operators(kinds: list[str], refs: list[int]) -> list[list[str]]:
    """The footer operator lists: a map source repeats DICT + SUB_EXPRESSION once per target it took
    (src/wizard/wizard.cpp:136-142 appends on every match)."""
    ops = []
    for c, k in enumerate(kinds):
        if k == "CONSTANT":
            ops.append(["CONSTANT"])
        elif k == "EQUAL":
            ops.append(["EQUAL"])
        elif k == "MAP_SOURCE":
            n_t = sum(1 for t, kt in enumerate(kinds) if kt == "MAP_TARGET" and refs[t] == c)
            ops.append(["DICT", "SUB_EXPRESSION"] * n_t)
        elif k == "MAP_TARGET":
            ops.append(["FIXME"])
        else:
            ops.append(["ZERO_COPY"])
    return ops


def spell_device(descs: list[tuple[int, int, int]], n_rows: int, stream=None) -> tuple[list[str], list[int]]:
    """descs: per column (data_t code, device values pointer, device offsets pointer or 0)."""
    n = len(descs)
    arr = (SpellColumn * max(n, 1))()
    for c, (t, v, o) in enumerate(descs):
        arr[c].type, arr[c].values, arr[c].offsets = t, v or None, o or None
    kinds = (C.c_int32 * max(n, 1))()
    refs = (C.c_int32 * max(n, 1))()
    fa.check(_lib().fls_wizard_spell(n, arr, n_rows, kinds, refs, fa.stream_ptr(stream)), "wizard_spell")
    return [KINDS[kinds[c]] for c in range(n)], [int(refs[c]) for c in range(n)]


def spell(columns, device=None, stream=None) -> list[dict]:
    """Plan of host columns (uploaded to the GPU for the analysis)."""
    import torch

    dev = device or torch.device("cuda:0")
    keep, descs, n_rows = [], [], None
    for col in columns:
        if isinstance(col, tuple):
            offs, data = col
            n = len(offs) - 1
            d_off = torch.from_numpy(np.ascontiguousarray(offs, dtype=np.uint64).view(np.int64)).to(dev)
            d_dat = torch.from_numpy(np.concatenate([np.asarray(data, np.uint8), np.zeros(8, np.uint8)])).to(dev)
            keep += [d_off, d_dat]
            descs.append((_STR, d_dat.data_ptr(), d_off.data_ptr()))
        else:
            a = np.ascontiguousarray(col)
            n = a.size
            d = torch.from_numpy(np.concatenate([a.view(np.uint8).ravel(), np.zeros(8, np.uint8)])).to(dev)
            keep.append(d)
            descs.append((fa.dtype_code(a.dtype), d.data_ptr(), 0))
        if n_rows is not None and n != n_rows:
            raise ValueError("columns of a row-group have equal length")
        n_rows = n
    kinds, refs = spell_device(descs, n_rows or 0, stream)
    torch.cuda.synchronize(dev)
    return [{"kind": k, "ref": r, "operators": o} for k, r, o in zip(kinds, refs, operators(kinds, refs))]
