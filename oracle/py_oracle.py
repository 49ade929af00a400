"""ctypes bindings for the CPU oracle (liboracle.so) and the compiled reference (oracle/_ref/libflsref.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
cpu_baseline leg. The shipped decoder never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libflsref.so")

# data_t (include/fastlanes.h)
DTYPE_CODE = {
    "float64": 1, "int8": 2, "int16": 3, "int32": 4, "int64": 5,
    "uint8": 6, "uint16": 7, "uint32": 8, "uint64": 9,
}
UNSIGNED = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}

_vp = C.c_void_p
_u8p = C.POINTER(C.c_uint8)


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class Oracle:
    """The repository's own C restatement (oracle/fls_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        lib = C.CDLL(path)
        self.lib = lib
        for name in ("orc_transpose", "orc_untranspose", "orc_unrsum"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp]
        for name in ("orc_pack", "orc_unpack"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp, C.c_uint]
        for name in ("orc_ffor", "orc_unffor"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp, C.c_uint, _vp]
        lib.orc_rsum.argtypes = [C.c_uint, _vp, _vp, _vp]
        lib.orc_ffor_analyze.argtypes = [C.c_uint, _vp, _vp]
        lib.orc_ffor_analyze.restype = C.c_uint8
        for name in ("orc_ffor_stream", "orc_unffor_stream"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp, C.c_uint64, C.c_uint, _vp]
        lib.orc_encode_column.argtypes = [C.c_uint8, C.c_uint8, _vp, C.c_uint64, _vp, C.c_uint64]
        lib.orc_encode_column.restype = C.c_int64
        lib.orc_encoded_size.argtypes = [C.c_uint8, C.c_uint8, _vp, C.c_uint64]
        lib.orc_encoded_size.restype = C.c_int64
        lib.orc_decode_column.argtypes = [C.c_uint8, _vp, C.c_uint64, _vp, C.c_uint64]
        lib.orc_decode_column.restype = C.c_int64
        lib.orc_encode_str_column.argtypes = [C.c_uint8, _vp, _vp, C.c_uint64, _vp, C.c_uint64]
        lib.orc_encode_str_column.restype = C.c_int64
        lib.orc_decode_str_column.argtypes = [_vp, C.c_uint64, _vp, _vp, C.c_uint64, C.POINTER(C.c_uint64)]
        lib.orc_decode_str_column.restype = C.c_int64
        lib.orc_hzl_encode.argtypes = [C.c_uint, _vp, C.c_uint64, _vp]
        lib.orc_hzl_encode.restype = C.c_int64
        lib.orc_hzl_decode.argtypes = [C.c_uint, _vp, _vp, C.POINTER(C.c_uint64)]
        lib.orc_hzl_decode.restype = C.c_int64
        lib.orc_fsst_build.argtypes = [_vp, _vp, C.c_uint64, _vp]
        lib.orc_fsst_build.restype = C.c_uint
        lib.orc_fsst_compress_one.argtypes = [_vp, _vp, C.c_size_t, _vp]
        lib.orc_fsst_compress_one.restype = C.c_size_t
        lib.orc_time_unffor_stream.argtypes = [C.c_uint, _vp, _vp, C.c_uint64, C.c_uint, _vp, C.c_int, C.c_int]
        lib.orc_time_unffor_stream.restype = C.c_double

    # -- primitives (one 1024-value vector) -------------------------------------------------------------------
    def transpose(self, x):
        out = np.empty_like(x)
        self.lib.orc_transpose(x.dtype.itemsize * 8, _ptr(x), _ptr(out))
        return out

    def untranspose(self, x):
        out = np.empty_like(x)
        self.lib.orc_untranspose(x.dtype.itemsize * 8, _ptr(x), _ptr(out))
        return out

    def ffor(self, x, bw, base):
        w = x.dtype.itemsize * 8
        out = np.zeros(max(bw, 1) * 128 // (w // 8), dtype=x.dtype)
        b = np.array([base], dtype=x.dtype)
        self.lib.orc_ffor(w, _ptr(x), _ptr(out), bw, _ptr(b))
        return out[: bw * 128 // (w // 8)]

    def unffor(self, packed, bw, base, dtype):
        w = np.dtype(dtype).itemsize * 8
        out = np.empty(1024, dtype=dtype)
        b = np.array([base], dtype=dtype)
        src = packed if packed.size else np.zeros(1, dtype=dtype)
        self.lib.orc_unffor(w, _ptr(src), _ptr(out), bw, _ptr(b))
        return out

    def pack(self, x, bw):
        w = x.dtype.itemsize * 8
        out = np.zeros(max(bw, 1) * 128 // (w // 8), dtype=x.dtype)
        self.lib.orc_pack(w, _ptr(x), _ptr(out), bw)
        return out[: bw * 128 // (w // 8)]

    def unpack(self, packed, bw, dtype):
        w = np.dtype(dtype).itemsize * 8
        out = np.empty(1024, dtype=dtype)
        src = packed if packed.size else np.zeros(1, dtype=dtype)
        self.lib.orc_unpack(w, _ptr(src), _ptr(out), bw)
        return out

    def rsum(self, x, base):
        out = np.empty_like(x)
        self.lib.orc_rsum(x.dtype.itemsize * 8, _ptr(x), _ptr(out), _ptr(base))
        return out

    def unrsum(self, x):
        out = np.empty_like(x)
        self.lib.orc_unrsum(x.dtype.itemsize * 8, _ptr(x), _ptr(out))
        return out

    def ffor_analyze(self, x):
        base = np.zeros(1, dtype=x.dtype)
        bw = self.lib.orc_ffor_analyze(x.dtype.itemsize * 8, _ptr(x), _ptr(base))
        return int(bw), base[0]

    # -- streams ----------------------------------------------------------------------------------------------
    def ffor_stream(self, x, bw, bases):
        w = x.dtype.itemsize * 8
        n_vec = x.size // 1024
        out = np.zeros(max(n_vec * bw * 128, 16), dtype=np.uint8)
        self.lib.orc_ffor_stream(w, _ptr(x), _ptr(out), n_vec, bw, _ptr(bases))
        return out[: n_vec * bw * 128]

    def unffor_stream(self, packed, n_vec, bw, bases, dtype):
        out = np.empty(n_vec * 1024, dtype=dtype)
        src = packed if packed.size else np.zeros(16, dtype=np.uint8)
        self.lib.orc_unffor_stream(np.dtype(dtype).itemsize * 8, _ptr(src), _ptr(out), n_vec, bw, _ptr(bases))
        return out

    # -- columns ----------------------------------------------------------------------------------------------
    def encode_column(self, x, expr_id):
        x = np.ascontiguousarray(x)
        dt = DTYPE_CODE[x.dtype.name]
        size = self.lib.orc_encoded_size(dt, expr_id, _ptr(x), x.size)
        if size < 0:
            raise ValueError(f"oracle cannot encode {x.dtype} with expression {expr_id} ({size})")
        out = np.empty(size, dtype=np.uint8)
        got = self.lib.orc_encode_column(dt, expr_id, _ptr(x), x.size, _ptr(out), size)
        assert got == size
        return out

    def decode_column(self, enc, dtype, n):
        out = np.empty(n, dtype=dtype)
        got = self.lib.orc_decode_column(DTYPE_CODE[np.dtype(dtype).name], _ptr(enc), enc.size, _ptr(out),
                                         out.nbytes)
        if got < 0:
            raise ValueError(f"oracle decode failed ({got})")
        return out[:got]

    def encode_str_column(self, offsets, data, expr_id=13):
        n = offsets.size - 1
        cap = int(offsets[-1]) * 3 + 64 * n + (1 << 20)
        out = np.empty(cap, dtype=np.uint8)
        got = self.lib.orc_encode_str_column(expr_id, _ptr(offsets), _ptr(data), n, _ptr(out), cap)
        if got < 0:
            raise ValueError(f"oracle cannot encode strings ({got})")
        return out[:got].copy()

    def decode_str_column(self, enc, n, max_bytes):
        offs = np.empty(n + 1, dtype=np.uint32)
        data = np.empty(max_bytes, dtype=np.uint8)
        nb = C.c_uint64(0)
        got = self.lib.orc_decode_str_column(_ptr(enc), enc.size, _ptr(offs), _ptr(data), max_bytes, C.byref(nb))
        if got < 0:
            raise ValueError(f"oracle string decode failed ({got})")
        return offs[: got + 1], data[: nb.value]

    def hzl_encode(self, x):
        w = x.dtype.itemsize * 8
        size = self.lib.orc_hzl_encode(w, _ptr(x), x.size, None)
        out = np.empty(size, dtype=np.uint8)
        self.lib.orc_hzl_encode(w, _ptr(x), x.size, _ptr(out))
        return out

    def hzl_decode(self, enc, n, dtype):
        out = np.empty(n + 1024, dtype=dtype)
        consumed = C.c_uint64(0)
        got = self.lib.orc_hzl_decode(np.dtype(dtype).itemsize * 8, _ptr(enc), _ptr(out), C.byref(consumed))
        return out[:got], consumed.value


class Reference:
    """The reference's own primitives and FSST, compiled from its sources by `make -C oracle ref`."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        self.lib = lib
        for name in ("ref_ffor", "ref_unffor"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp, C.c_uint, _vp]
        for name in ("ref_pack", "ref_unpack"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp, C.c_uint]
        lib.ref_rsum.argtypes = [C.c_uint, _vp, _vp, _vp]
        for name in ("ref_unrsum", "ref_transpose_i", "ref_untranspose_i", "ref_untranspose_o"):
            getattr(lib, name).argtypes = [C.c_uint, _vp, _vp]
        lib.ref_fsst_encode.argtypes = [_vp, _vp, C.c_uint64, _vp, _vp, C.c_uint64, _vp]
        lib.ref_fsst_encode.restype = C.c_uint
        lib.ref_fsst_decode.argtypes = [_vp, _vp, _vp, C.c_uint64, _vp, C.c_uint64, _vp]
        lib.ref_fsst_decode.restype = C.c_uint64
        lib.ref_time_unffor.argtypes = [C.c_uint, _vp, _vp, C.c_uint64, C.c_uint, _vp, C.c_int, C.c_int, C.c_int]
        lib.ref_time_unffor.restype = C.c_double

    def ffor(self, x, bw, base):
        w = x.dtype.itemsize * 8
        out = np.zeros(max(bw, 1) * 128 // (w // 8) + 16, dtype=x.dtype)
        b = np.array([base], dtype=x.dtype)
        self.lib.ref_ffor(w, _ptr(x), _ptr(out), bw, _ptr(b))
        return out[: bw * 128 // (w // 8)]

    def unffor(self, packed, bw, base, dtype):
        w = np.dtype(dtype).itemsize * 8
        out = np.empty(1024, dtype=dtype)
        b = np.array([base], dtype=dtype)
        src = np.zeros(max(packed.size, 1) + 16, dtype=dtype)
        src[: packed.size] = packed
        self.lib.ref_unffor(w, _ptr(src), _ptr(out), bw, _ptr(b))
        return out

    def pack(self, x, bw):
        w = x.dtype.itemsize * 8
        out = np.zeros(max(bw, 1) * 128 // (w // 8) + 16, dtype=x.dtype)
        self.lib.ref_pack(w, _ptr(x), _ptr(out), bw)
        return out[: bw * 128 // (w // 8)]

    def unpack(self, packed, bw, dtype):
        w = np.dtype(dtype).itemsize * 8
        out = np.empty(1024, dtype=dtype)
        src = np.zeros(max(packed.size, 1) + 16, dtype=dtype)
        src[: packed.size] = packed
        self.lib.ref_unpack(w, _ptr(src), _ptr(out), bw)
        return out

    def rsum(self, x, base):
        out = np.empty_like(x)
        self.lib.ref_rsum(x.dtype.itemsize * 8, _ptr(x), _ptr(out), _ptr(base))
        return out

    def unrsum(self, x):
        src = np.zeros(2048, dtype=x.dtype)  # the generated unrsum reads one row past the vector
        src[:1024] = x
        out = np.empty(2048, dtype=x.dtype)
        self.lib.ref_unrsum(x.dtype.itemsize * 8, _ptr(src), _ptr(out))
        return out[:1024]

    def transpose(self, x):
        out = np.empty_like(x)
        self.lib.ref_transpose_i(x.dtype.itemsize * 8, _ptr(x), _ptr(out))
        return out

    def untranspose(self, x):
        out = np.empty_like(x)
        self.lib.ref_untranspose_i(x.dtype.itemsize * 8, _ptr(x), _ptr(out))
        return out

    def untranspose_o(self, x):
        out = np.empty_like(x)
        self.lib.ref_untranspose_o(x.dtype.itemsize * 8, _ptr(x), _ptr(out))
        return out

    def fsst_encode(self, offsets, data):
        n = offsets.size - 1
        table = np.zeros(4096, dtype=np.uint8)
        cap = int(offsets[-1]) * 2 + 8 * n + 4096
        comp = np.zeros(cap, dtype=np.uint8)
        coff = np.zeros(n + 1, dtype=np.uint32)
        tl = self.lib.ref_fsst_encode(_ptr(offsets), _ptr(data), n, _ptr(table), _ptr(comp), cap, _ptr(coff))
        if tl == 0:
            raise RuntimeError("reference fsst_compress did not fit")
        return table[:tl].copy(), comp[: coff[-1]].copy(), coff

    def fsst_decode(self, table, comp, coff, max_bytes):
        n = coff.size - 1
        out = np.zeros(max_bytes + 64, dtype=np.uint8)
        offs = np.zeros(n + 1, dtype=np.uint32)
        got = self.lib.ref_fsst_decode(_ptr(table), _ptr(comp), _ptr(coff), n, _ptr(out), max_bytes + 64,
                                       _ptr(offs))
        return offs, out[:got].copy()


def reference_available() -> bool:
    return os.path.exists(REF_SO)
