"""fastlanes_amd -- FastLanes decoding on MI355X (gfx950).

Python access to the C ABI of include/fastlanes.h and include/fls_gpu.h (libfastlanes_amd.so, built in-tree by
`make -C fastlanes_amd`). Device buffers are passed as raw pointers (torch CUDA tensors' data_ptr() work).
Importing this module does not touch the GPU; loading the library fails loudly if it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastlanes_amd.so")

# data_t (include/fastlanes.h)
DATA_T = {"float64": 1, "int8": 2, "int16": 3, "int32": 4, "int64": 5, "uint8": 6, "uint16": 7, "uint32": 8,
          "uint64": 9, "str": 10}
FLS_OK, FLS_E_DATA = 0, -6

_lib = None
_vp, _u64, _u32, _int = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int


class FlsError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """The loaded C ABI. Raises if the HIP library is missing (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlsError(f"{LIB_PATH} is not built: run `make -C fastlanes_amd` (hipcc --offload-arch=gfx950)")
    # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7 and loads it by path. Loaded first, it is
    # the runtime this library binds to (same soname); loaded after this library, it would be a second runtime
    # beside the system one, and the one initialised later finds no device. So PyTorch, when installed, goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.fls_gpu_error_string.restype = C.c_char_p
    L.fls_gpu_plan_create.argtypes = [_vp, _u64, C.c_uint8, C.POINTER(_vp)]
    L.fls_gpu_plan_destroy.argtypes = [_vp]
    L.fls_gpu_plan_destroy.restype = None
    for n in ("fls_gpu_plan_n_values", "fls_gpu_plan_encoded_size"):
        getattr(L, n).argtypes = [_vp]
        getattr(L, n).restype = _u64
    L.fls_gpu_plan_n_pages.argtypes = [_vp]
    L.fls_gpu_plan_n_pages.restype = _u32
    L.fls_gpu_plan_decode.argtypes = [_vp, _vp, _vp, _u64, _vp]
    L.fls_gpu_plan_check.argtypes = [_vp, _vp]
    L.fls_gpu_plan_select.argtypes = [_vp, _vp, C.c_int, _u64, _vp, _vp, _vp]
    L.fls_gpu_selection_indices.argtypes = [_vp, _u64, _vp, _vp, _vp]
    L.fls_gpu_selection_and.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp]
    L.fls_gpu_encode.argtypes = [_vp, _u64, C.c_uint8, C.c_uint8, _vp, _u64, C.POINTER(_u64), _vp]
    L.fls_gpu_encode_noalloc.argtypes = [_vp, _u64, C.c_uint8, C.c_uint8, _vp, _u64, C.POINTER(_u64), _vp]
    L.fls_gpu_plan_string_offsets.argtypes = [_vp, _vp, _vp, _vp]
    L.fls_gpu_plan_decode_strings.argtypes = [_vp, _vp, _vp, _vp, _u64, _vp]
    L.fls_gpu_plan_decode_strings_fused.argtypes = [_vp, _vp, _vp, _vp, _u64, _vp]
    L.fls_gpu_unffor_stream.argtypes = [C.c_uint, C.c_uint, _vp, _vp, _vp, _u64, _vp]
    L.fls_gpu_rowgroup_create.argtypes = [_u32, C.POINTER(_vp), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(C.c_uint8),
                                          C.POINTER(_vp)]
    L.fls_gpu_rowgroup_destroy.argtypes = [_vp]
    L.fls_gpu_rowgroup_destroy.restype = None
    L.fls_gpu_rowgroup_layout.argtypes = [_vp, C.POINTER(_u64), C.POINTER(_u64)]
    L.fls_gpu_rowgroup_decode.argtypes = [_vp, _vp, _vp, _int, _vp]
    L.fls_gpu_rowgroup_decode_graph.argtypes = [_vp, _vp, _vp, _vp]
    L.fls_gpu_rowgroup_check.argtypes = [_vp, _vp]
    L.fls_gpu_tune_set.argtypes = [_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
    L.fls_gpu_tune_get.argtypes = [_int, C.c_uint, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                   C.POINTER(C.c_uint)]
    L.fls_gpu_rowgroup_work_list.argtypes = [_vp, C.POINTER(_u32), _vp, _u32, C.POINTER(_u32), C.POINTER(_u32),
                                             C.POINTER(_u32), _vp]
    L.fls_rowgroup_plan_work_list.argtypes = [_u32, C.POINTER(_vp), C.POINTER(_u64), C.POINTER(_u64),
                                              C.POINTER(C.c_uint8), _u32, C.POINTER(_u32), _vp, _u32, _vp,
                                              C.POINTER(_u32)]
    # include/fastlanes.h
    L.fls_cpi_connection_open.restype = _vp
    L.fls_cpi_connection_close.argtypes = [_vp]
    L.fls_cpi_connection_close.restype = None
    L.encode_from_memory.argtypes = [_vp, _u64, _u64, _vp, C.POINTER(_u64), C.POINTER(_u64), C.c_uint8, C.c_uint8]
    L.encode_from_memory.restype = C.c_bool
    L.decode_to_memory.argtypes = [_vp, _vp, C.c_uint8]
    L.decode_to_memory.restype = None
    L.init_random_access.argtypes = [_vp, _u64, C.POINTER(_u64), C.POINTER(_u64)]
    L.init_random_access.restype = None
    L.random_access.argtypes = [_vp, _u64, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_vp), _vp]
    L.random_access.restype = None
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != FLS_OK:
        raise FlsError(f"{what}: {rc} {lib().fls_gpu_error_string().decode()}")


def dtype_code(dtype) -> int:
    return DATA_T[np.dtype(dtype).name] if dtype != "str" else DATA_T["str"]


def stream_ptr(stream=None) -> int:
    """hipStream_t of a torch stream (None -> torch's current stream)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


class Plan:
    """Host-side parse of one encoded column; decodes device-resident copies of the same bytes."""

    def __init__(self, host_encoded: np.ndarray, dtype):
        host_encoded = np.ascontiguousarray(host_encoded, dtype=np.uint8)
        self.dtype = dtype
        h = _vp()
        check(lib().fls_gpu_plan_create(host_encoded.ctypes.data, host_encoded.size, dtype_code(dtype), C.byref(h)),
              "plan_create")
        self.h = h

    @property
    def n_values(self) -> int:
        return lib().fls_gpu_plan_n_values(self.h)

    @property
    def encoded_size(self) -> int:
        return lib().fls_gpu_plan_encoded_size(self.h)

    @property
    def n_pages(self) -> int:
        return lib().fls_gpu_plan_n_pages(self.h)

    def decode(self, dev_encoded_ptr: int, dev_out_ptr: int, out_capacity: int, stream: int = 0) -> None:
        check(lib().fls_gpu_plan_decode(self.h, dev_encoded_ptr, dev_out_ptr, out_capacity, stream), "plan_decode")

    def check(self, stream: int = 0) -> None:
        check(lib().fls_gpu_plan_check(self.h, stream), "plan_check")

    def select(self, dev_encoded_ptr: int, op: str, constant, dev_bitmap_ptr: int, dev_counts_ptr: int,
               stream: int = 0) -> None:
        """Predicate pushdown: per vector, 16 bitmap words of `value <op> constant` and the selected count
        (op in ==, >, <, >=, <=, !=); no decoded values are written."""
        bits = np.array([constant], dtype=np.dtype(self.dtype)).view(np.uint8)
        c = int.from_bytes(bits.tobytes(), "little")
        check(lib().fls_gpu_plan_select(self.h, dev_encoded_ptr, PRED_OPS[op], c, dev_bitmap_ptr, dev_counts_ptr,
                                        stream), "plan_select")

    def string_offsets(self, dev_encoded_ptr: int, dev_offsets_ptr: int, stream: int = 0) -> None:
        check(lib().fls_gpu_plan_string_offsets(self.h, dev_encoded_ptr, dev_offsets_ptr, stream), "string_offsets")

    def decode_strings(self, dev_encoded_ptr, dev_offsets_ptr, dev_bytes_ptr, cap, stream: int = 0) -> None:
        check(lib().fls_gpu_plan_decode_strings(self.h, dev_encoded_ptr, dev_offsets_ptr, dev_bytes_ptr, cap, stream),
              "decode_strings")

    def decode_strings_fused(self, dev_encoded_ptr, dev_offsets_ptr, dev_bytes_ptr, cap, stream: int = 0) -> None:
        """Offsets and bytes in one pass (one launch per page)."""
        check(lib().fls_gpu_plan_decode_strings_fused(self.h, dev_encoded_ptr, dev_offsets_ptr, dev_bytes_ptr, cap,
                                                      stream), "decode_strings_fused")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fls_gpu_plan_destroy(self.h)
            self.h = None


def gpu_encode(dev_values: int, n: int, dtype, expr_id: int, dev_out: int, capacity: int, stream: int = 0) -> int:
    """Encodes n device-resident values on the GPU (expressions 0, 1, 11, 6, 7); returns the encoded size."""
    size = _u64(0)
    check(lib().fls_gpu_encode(dev_values, n, dtype_code(dtype), expr_id, dev_out, capacity, C.byref(size), stream),
          "gpu_encode")
    return size.value


PRED_OPS = {"==": 0, ">": 1, "<": 2, ">=": 3, "<=": 4, "!=": 5}


def selection_indices(dev_bitmap: int, n_vec: int, dev_indices: int, dev_counts: int | None = None,
                      stream: int = 0) -> None:
    """Index arrays (uint16, 1024 slots per vector) of the selected positions, ascending."""
    check(lib().fls_gpu_selection_indices(dev_bitmap, n_vec, dev_indices, dev_counts, stream), "selection_indices")


def selection_and(dev_a: int, dev_b: int, n_vec: int, dev_out: int, dev_counts: int | None = None,
                  stream: int = 0) -> None:
    check(lib().fls_gpu_selection_and(dev_a, dev_b, n_vec, dev_out, dev_counts, stream), "selection_and")


# One tile of the persistent row-group launch (struct RgTile, fastlanes_amd/csrc/rowgroup.hpp): vectors
# [v0, v0 + nv) of the page at byte `page` of the encoded buffer, decoded to byte `out` of the output buffer; its
# input is `n_rng` byte ranges of the page (rsrc, rlen) staged at rlds in the wave's LDS slot.
RG_TILE = np.dtype([("page", "<u8"), ("out", "<u8"), ("dict", "<u8"), ("v0", "<u4"), ("dict_n", "<u4"),
                    ("vec_c", "<u4"), ("nv", "<u2"), ("kind", "<u2"), ("arr_c", "<u2"), ("n_rng", "<u2"),
                    ("off", "<u4", (5,)), ("rsrc", "<u4", (3,)), ("rlen", "<u4", (3,)), ("rlds", "<u4", (3,)),
                    ("chunks", "<u4"), ("sdelta", "<i4", (3,)), ("ldelta", "<i4", (3,))])
assert RG_TILE.itemsize == 128
# Tile kinds (enum RgKind): 4 element widths each of UNCOMPRESSED, FFOR, DELTA, RLE and dictionary codes
# (UNCOMPRESSED, FFOR, DELTA).
RG_KIND_NAMES = [f"{k}{w}" for k in ("unc", "ffor", "delta", "rle", "dict1_", "dict2_", "dict7_") for w in (8, 16, 32, 64)]


def plan_work_list(host_columns, dtypes, device_offsets, waves: int = 0):
    """The persistent row-group launch's work list planned on the host alone (fls_rowgroup_plan_work_list):
    (tiles, LDS bytes per wave, wave runs). Dictionary tiles carry a placeholder instead of a device address. With
    waves > 0 the tiles come in launch order and wave w's run is tiles[runs[w]:runs[w + 1]]."""
    n = len(host_columns)
    cols = [np.ascontiguousarray(c, dtype=np.uint8) for c in host_columns]
    ptrs = (_vp * max(n, 1))(*[c.ctypes.data for c in cols])
    sizes = (_u64 * max(n, 1))(*[c.size for c in cols])
    offs = (_u64 * max(n, 1))(*[int(o) for o in device_offsets])
    types = (C.c_uint8 * max(n, 1))(*[dtype_code(d) for d in dtypes])
    cnt, wlds = _u32(0), _u32(0)
    check(lib().fls_rowgroup_plan_work_list(n, ptrs, sizes, offs, types, 0, C.byref(cnt), None, 0, None,
                                            C.byref(wlds)), "plan_work_list")
    tiles = np.zeros(cnt.value, dtype=RG_TILE)
    runs = np.zeros(waves + 1, dtype=np.uint32)
    check(lib().fls_rowgroup_plan_work_list(n, ptrs, sizes, offs, types, waves, C.byref(cnt), tiles.ctypes.data,
                                            cnt.value, runs.ctypes.data if waves else None, C.byref(wlds)),
          "plan_work_list")
    return tiles, int(wlds.value), runs


class RowGroup:
    """Many encoded columns decoded together (fls_gpu_rowgroup_*). The encoded columns live in one device buffer at
    `device_offsets`; decoded columns land in one output buffer at `offsets` (256-byte aligned)."""

    def __init__(self, host_columns, dtypes, device_offsets):
        n = len(host_columns)
        self._cols = [np.ascontiguousarray(c, dtype=np.uint8) for c in host_columns]
        ptrs = (_vp * max(n, 1))(*[c.ctypes.data for c in self._cols])
        sizes = (_u64 * max(n, 1))(*[c.size for c in self._cols])
        offs = (_u64 * max(n, 1))(*[int(o) for o in device_offsets])
        types = (C.c_uint8 * max(n, 1))(*[dtype_code(d) for d in dtypes])
        h = _vp()
        check(lib().fls_gpu_rowgroup_create(n, ptrs, sizes, offs, types, C.byref(h)), "rowgroup_create")
        self.h = h
        self.n = n
        out_offs = (_u64 * max(n, 1))()
        total = _u64(0)
        check(lib().fls_gpu_rowgroup_layout(h, out_offs, C.byref(total)), "rowgroup_layout")
        self.offsets = [int(out_offs[i]) for i in range(n)]
        self.total_bytes = int(total.value)

    def decode(self, dev_encoded: int, dev_out: int, mode: int = 0, stream: int = 0) -> None:
        """mode 0: one persistent launch over the work list of (page, vector range) tiles; mode 1: one launch per page
        in column order; mode 2: the same launches over 4 parallel stream branches."""
        check(lib().fls_gpu_rowgroup_decode(self.h, dev_encoded, dev_out, mode, stream), "rowgroup_decode")

    def decode_graph(self, dev_encoded: int, dev_out: int, stream: int = 0) -> None:
        """The persistent decode (mode 0) captured once in a HIP graph, then replayed."""
        check(lib().fls_gpu_rowgroup_decode_graph(self.h, dev_encoded, dev_out, stream), "rowgroup_decode_graph")

    def work_list(self):
        """(tiles as a RG_TILE structured array in launch order, blocks, LDS bytes per wave, wave runs) of the
        persistent launch: wave w decodes tiles[runs[w]:runs[w + 1]]."""
        n, blocks, waves, wlds = _u32(0), _u32(0), _u32(0), _u32(0)
        check(lib().fls_gpu_rowgroup_work_list(self.h, C.byref(n), None, 0, C.byref(blocks), C.byref(waves), None,
                                               None), "rowgroup_work_list")
        tiles = np.zeros(n.value, dtype=RG_TILE)
        runs = np.zeros(waves.value * blocks.value + 1, dtype=np.uint32)
        check(lib().fls_gpu_rowgroup_work_list(self.h, C.byref(n), tiles.ctypes.data, n.value, C.byref(blocks),
                                               C.byref(waves), C.byref(wlds), runs.ctypes.data),
              "rowgroup_work_list")
        return tiles, int(blocks.value), int(wlds.value), runs

    def check(self, stream: int = 0) -> None:
        check(lib().fls_gpu_rowgroup_check(self.h, stream), "rowgroup_check")

    def decode_strings_fused(self, dev_encoded_ptr, dev_offsets_ptr, dev_bytes_ptr, cap, stream: int = 0) -> None:
        """Offsets and bytes in one pass (one launch per page)."""
        check(lib().fls_gpu_plan_decode_strings_fused(self.h, dev_encoded_ptr, dev_offsets_ptr, dev_bytes_ptr, cap,
                                                      stream), "decode_strings_fused")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fls_gpu_rowgroup_destroy(self.h)
            self.h = None


def unffor_stream(width: int, bw: int, dev_packed: int, dev_bases: int | None, dev_out: int, n_vec: int,
                  stream: int = 0) -> None:
    check(lib().fls_gpu_unffor_stream(width, bw, dev_packed, dev_bases or None, dev_out, n_vec, stream),
          "unffor_stream")


def encode_from_memory(values: np.ndarray, expr_id: int, capacity: int | None = None) -> np.ndarray:
    """include/fastlanes.h encode_from_memory on a host array; returns the encoded bytes."""
    values = np.ascontiguousarray(values)
    # default room: dictionary pages hold 4-byte codes whatever the value width, plus the dictionary itself
    cap = capacity if capacity is not None else values.size * 4 + values.nbytes * 2 + (1 << 16) + 64 * (values.size // 1024)
    out = np.empty(cap, dtype=np.uint8)
    size, meta = _u64(0), _u64(0)
    ok = lib().encode_from_memory(values.ctypes.data, values.size, cap, out.ctypes.data, C.byref(size), C.byref(meta),
                                  dtype_code(values.dtype), expr_id)
    if not ok:
        raise FlsError(f"encode_from_memory failed: {lib().fls_gpu_error_string().decode()}")
    return out[: size.value].copy()


def decode_to_memory(encoded_ptr: int, out_ptr: int, dtype) -> None:
    """include/fastlanes.h decode_to_memory (host or device pointers)."""
    lib().decode_to_memory(encoded_ptr, out_ptr, dtype_code(dtype))


def encode_strings(offsets: np.ndarray, data: np.ndarray) -> np.ndarray:
    """FSST-encode a string column given Arrow-style uint32 offsets (n+1) and bytes (fls_encode_strings)."""
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = offsets.size - 1
    cap = int(offsets[-1]) * 2 + 8 * n + (1 << 16)
    out = np.empty(cap, dtype=np.uint8)
    size = _u64(0)
    L = lib()
    L.fls_encode_strings.argtypes = [_vp, _vp, _u64, _vp, _u64, C.POINTER(_u64)]
    check(L.fls_encode_strings(offsets.ctypes.data, data.ctypes.data if data.size else None, n, out.ctypes.data, cap,
                               C.byref(size)), "encode_strings")
    return out[: size.value].copy()
