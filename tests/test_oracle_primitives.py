"""The CPU oracle's primitives against the reference's own golden vectors and compiled code."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

W2DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


@pytest.fixture(scope="module")
def kat():
    return np.load(os.path.join(GOLDEN, "rsum_kat.npz"))


@pytest.fixture(scope="module")
def helper():
    return np.load(os.path.join(GOLDEN, "unffor_helper.npz"))


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_rsum_known_answer(oracle, kat, W):
    # gen/rsum_test.cpp:40-79: rsum(delta, base) == transposed; untranspose_o(transposed) == untransposed
    delta, base = kat[f"delta_arr_ow{W}"], kat[f"base_arr_ow{W}"]
    assert base.size == 1024 // W
    t = oracle.rsum(delta, base)
    np.testing.assert_array_equal(t, kat[f"transposed_arr_ow{W}"])
    np.testing.assert_array_equal(oracle.untranspose(t), kat[f"untransposed_arr_ow{W}"])
    np.testing.assert_array_equal(oracle.transpose(kat[f"untransposed_arr_ow{W}"]), t)


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_unrsum_known_answer(oracle, kat, W):
    # gen/unrsum_test.cpp:49-100: unrsum(transposed) == delta (first step of each lane is 0)
    np.testing.assert_array_equal(oracle.unrsum(kat[f"transposed_arr_ow{W}"]), kat[f"delta_arr_ow{W}"])


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_transpose_is_a_permutation_and_inverse(oracle, W):
    x = np.arange(1024, dtype=np.uint64).astype(W2DT[W]) if W > 8 else (np.arange(1024) % 256).astype(np.uint8)
    perm = oracle.untranspose(np.arange(1024, dtype=np.uint16)) if W == 16 else None
    if perm is not None:
        assert sorted(perm.tolist()) == list(range(1024))
    np.testing.assert_array_equal(oracle.untranspose(oracle.transpose(x)), x)


def test_helper_arrays_ffor_match_reference_bytes(oracle, helper):
    # the 124 seeded arrays of gen/unffor_helper.hpp, ffor'd by the compiled reference with base 10
    n = 0
    for key in helper.files:
        if not key.startswith("in_"):
            continue
        _, b, w = key.split("_")
        bw, W = int(b[1:]), int(w[1:])
        x = helper[key]
        packed = oracle.ffor(x, bw, W2DT[W](10))
        np.testing.assert_array_equal(packed, helper[f"ffor10_b{bw}_w{W}"], err_msg=key)
        np.testing.assert_array_equal(oracle.unffor(packed, bw, W2DT[W](10), W2DT[W]), x, err_msg=key)
        p2 = oracle.pack(x, bw)
        np.testing.assert_array_equal(p2, helper[f"pack_b{bw}_w{W}"], err_msg=key)
        mask = (1 << bw) - 1 if bw < 64 else (1 << 64) - 1
        np.testing.assert_array_equal(oracle.unpack(p2, bw, W2DT[W]), (x.astype(np.uint64) & np.uint64(mask)).astype(W2DT[W]))
        n += 1
    assert n == 124


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_primitives_match_compiled_reference(oracle, reference, W):
    rng = np.random.default_rng(W)
    dt = W2DT[W]
    for bw in range(W + 1):
        x = rng.integers(0, np.iinfo(dt).max, 1024, dtype=dt, endpoint=True)
        base = dt(rng.integers(0, np.iinfo(dt).max, dtype=dt, endpoint=True))
        a = oracle.ffor(x, bw, base)
        np.testing.assert_array_equal(a, reference.ffor(x, bw, base))
        np.testing.assert_array_equal(oracle.unffor(a, bw, base, dt), reference.unffor(a, bw, base, dt))
        np.testing.assert_array_equal(oracle.pack(x, bw), reference.pack(x, bw))
    x = rng.integers(0, np.iinfo(dt).max, 1024, dtype=dt, endpoint=True)
    np.testing.assert_array_equal(oracle.transpose(x), reference.transpose(x))
    np.testing.assert_array_equal(oracle.untranspose(x), reference.untranspose(x))
    np.testing.assert_array_equal(oracle.untranspose(x), reference.untranspose_o(x))
    np.testing.assert_array_equal(oracle.unrsum(x), reference.unrsum(x))
    np.testing.assert_array_equal(oracle.rsum(x, x[: 1024 // W].copy()), reference.rsum(x, x[: 1024 // W].copy()))


@pytest.mark.parametrize("W", [8, 16, 32, 64])
def test_ffor_special_widths(oracle, W):
    dt = W2DT[W]
    x = np.full(1024, 77, dtype=dt)
    assert oracle.ffor(x, 0, dt(77)).size == 0  # bw 0 writes nothing
    np.testing.assert_array_equal(oracle.unffor(np.zeros(0, dtype=dt), 0, dt(77), dt), x)
    y = np.random.default_rng(1).integers(0, np.iinfo(dt).max, 1024, dtype=dt, endpoint=True)
    packed = oracle.ffor(y, W, dt(5))  # bw == W: value - base stored whole, decode wraps mod 2^W
    np.testing.assert_array_equal(packed, oracle.transpose(oracle.untranspose((y - dt(5)).astype(dt))))
    np.testing.assert_array_equal(oracle.unffor(packed, W, dt(5), dt), y)


def test_ffor_analyze_bit_width(oracle):
    lib = oracle.lib
    lib.orc_ffor_bw_reference_formula.restype = __import__("ctypes").c_uint8
    lib.orc_ffor_bw_reference_formula.argtypes = [__import__("ctypes").c_uint64]
    for W in (8, 16, 32, 64):
        dt = W2DT[W]
        for k in range(W + 1):
            for delta in {max((1 << k) - 1, 0), 1 << k if k < W else (1 << W) - 1}:
                x = np.zeros(1024, dtype=dt)
                x[5] = dt(delta) if delta < (1 << W) else dt((1 << W) - 1)
                bw, base = oracle.ffor_analyze(x)
                d = int(x.max())
                assert bw == d.bit_length(), (W, k, d)
                assert base == 0
                if d + 1 < (1 << 40):  # where doubles hold every integer the two formulas must agree
                    assert bw == lib.orc_ffor_bw_reference_formula(d)
    # the reference's float formula drops a bit at delta = 2^50 (decode would be wrong): documented deviation
    assert lib.orc_ffor_bw_reference_formula(1 << 50) == 50 and (1 << 50).bit_length() == 51
