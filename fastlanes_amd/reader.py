"""Connection / Reader mirror for the decode path: a row-group written to a directory and read back through one
device-side decode.

Follows the reference's file path (`Connection::to_fls` / `Connection::read_fls`, src/connection.cpp:102-128;
`Reader(dir, Connection&)`, `Reader::get_chunk(vec_idx)`, src/reader/reader.cpp:19-61): a directory holds
`data.fls` (the encoded columns back to back, 256-byte aligned) and `footer.json` in the reference's keys
(src/json/fls_json.cpp:30-61; include/fls/footer/rowgroup_footer.hpp:23-75): "1: N VEC", "2: Rowgroup size",
"3: Column Metadatas" with per column "logical_type", "logical_type_str", "Encoding RPN", "col_ofs" and
"is_determined", plus "size" (the reference's ColumnMetadata::size, which its JSON leaves out) and "pool id" (the
expression). The C++ Reader of include/fls_reader.hpp reads the same directories. Differences: "Encoding RPN" names
the column's page expression ([ExpT name, data type name], as the reference flattens its operators), and
`get_chunk` returns device tensors of the whole row-group decoded by one row-group launch (replayed from a HIP
graph) rather than decoding vector by vector on the host.
"""
from __future__ import annotations

import json
import os

import numpy as np

import fastlanes_amd as fa

FASTLANES_FILE_NAME = "data.fls"
FOOTER_FILE_NAME = "footer.json"
N_VEC, ROWGROUP_SIZE, COLUMN_MTDS = "1: N VEC", "2: Rowgroup size", "3: Column Metadatas"
_TORCH_DT = {"int8": "int8", "int16": "int16", "int32": "int32", "int64": "int64", "uint8": "uint8",
             "uint16": "uint16", "uint32": "uint32", "uint64": "uint64", "float64": "float64"}
# reference include/fls/expression/data_type.hpp:9-29
DATA_TYPES = {"float64": (1, "DOUBLE"), "int8": (2, "INT8"), "int16": (3, "INT16"), "int32": (4, "INT32"),
              "int64": (5, "INT64"), "uint8": (6, "UINT8"), "uint16": (7, "UINT16"), "uint32": (8, "UINT32"),
              "uint64": (9, "UINT64")}
_NUMPY_OF = {code: np.dtype(name) for name, (code, _) in DATA_TYPES.items()}


def exp_name(pool_id: int) -> str:
    """ExpT of an integer-pool expression (reference src/cor/exp_pool.cpp:23-65, include/fls/cor/exp/exp_type.hpp)."""
    if pool_id == 0:
        return "UNCOMPRESSED"
    for hi, name in ((5, "FFOR"), (10, "RLE"), (15, "DELTA"), (26, "DICT"), (27, "DICT_RLE"), (38, "DICT_FRQ")):
        if pool_id <= hi:
            return name
    raise ValueError(f"pool id {pool_id}")


def column_dtypes(footer: dict) -> list:
    return [_NUMPY_OF[int(m["logical_type"])] for m in footer[COLUMN_MTDS]]


def column_sizes(footer: dict) -> list:
    """Encoded bytes per column: "size", or up to the next column's offset (the end of the row-group)."""
    metas = footer[COLUMN_MTDS]
    out = []
    for m in metas:
        if "size" in m:
            out.append(int(m["size"]))
            continue
        ends = [int(o["col_ofs"]) for o in metas if int(o["col_ofs"]) > int(m["col_ofs"])]
        out.append(min(ends, default=int(footer[ROWGROUP_SIZE])) - int(m["col_ofs"]))
    return out


class Connection:
    """Entry points of fastlanes::Connection used on the decode path (include/fls/connection.hpp:65-92)."""

    encode_from_memory = staticmethod(fa.encode_from_memory)
    decode_to_memory = staticmethod(fa.decode_to_memory)

    @staticmethod
    def to_fls(dir_path: str, columns, expr_ids) -> dict:
        """Encode equal-length columns with the given expression ids into dir_path/{data.fls, footer.json}."""
        os.makedirs(dir_path, exist_ok=True)
        if len(columns) != len(expr_ids):
            raise ValueError("one expression id per column")
        n = {int(np.asarray(c).size) for c in columns}
        if len(n) > 1:
            raise ValueError("columns of a row-group have equal length")
        metas, pos = [], 0
        with open(os.path.join(dir_path, FASTLANES_FILE_NAME), "wb") as f:
            for c, eid in zip(columns, expr_ids):
                c = np.ascontiguousarray(c)
                enc = fa.encode_from_memory(c, int(eid))
                pad = (-pos) % 256  # columns start 256-byte aligned in the file (and on the device)
                f.write(b"\0" * pad)
                pos += pad
                code, name = DATA_TYPES[c.dtype.name]
                metas.append({"logical_type": code, "logical_type_str": name,
                              "Encoding RPN": {"operators": [exp_name(int(eid)), name], "operands": [len(metas)]},
                              "col_ofs": pos, "is_determined": True, "size": int(enc.size), "pool id": int(eid)})
                f.write(enc.tobytes())
                pos += enc.size
        footer = {N_VEC: (n.pop() if n else 0) // 1024, ROWGROUP_SIZE: pos, COLUMN_MTDS: metas}
        with open(os.path.join(dir_path, FOOTER_FILE_NAME), "w") as f:
            json.dump(footer, f, indent=1)
        return footer

    def read_fls(self, dir_path: str) -> "Reader":
        return Reader(dir_path, self)

    @staticmethod
    def spell(columns) -> list[dict]:
        """Connection::spell (src/connection.cpp:91-100) -> Wizard::Spell: the structural plan of a row-group's columns
        (CONSTANT / EQUAL / one-to-one map / ZERO_COPY per column), computed on the GPU (fastlanes_amd.spell)."""
        from fastlanes_amd import spell as sp

        return sp.spell(columns)


class Chunk:
    """One vector index of every column: device tensors of (up to) 1024 values each."""

    def __init__(self, columns):
        self.columns = columns

    def __len__(self):
        return len(self.columns)

    def __getitem__(self, i):
        return self.columns[i]


class Reader:
    def __init__(self, dir_path: str, connection: Connection | None = None, device=None, stream=None):
        import torch

        with open(os.path.join(dir_path, FOOTER_FILE_NAME)) as f:
            self.footer = json.load(f)
        data = np.fromfile(os.path.join(dir_path, FASTLANES_FILE_NAME), dtype=np.uint8)
        if data.size != self.footer[ROWGROUP_SIZE]:
            raise fa.FlsError(f"{dir_path}: data size {data.size} != footer \"{ROWGROUP_SIZE}\"")
        self.device = device or torch.device("cuda:0")
        self.stream = stream
        metas = self.footer[COLUMN_MTDS]
        self.dtypes = column_dtypes(self.footer)
        offs = [int(m["col_ofs"]) for m in metas]
        cols = [data[o: o + sz] for o, sz in zip(offs, column_sizes(self.footer))]
        self.rg = fa.RowGroup(cols, self.dtypes, offs)
        self.d_enc = torch.from_numpy(np.concatenate([data, np.zeros(16, np.uint8)])).to(self.device)
        self.d_out = torch.empty(self.rg.total_bytes + 256, dtype=torch.uint8, device=self.device)
        self._decoded = False
        self.n_vec = int(self.footer[N_VEC])

    def _decode(self):
        sp = fa.stream_ptr(self.stream)
        self.rg.decode_graph(self.d_enc.data_ptr(), self.d_out.data_ptr(), sp)
        self.rg.check(sp)
        self._decoded = True

    def get_chunk(self, vec_idx: int) -> Chunk:
        import torch

        if not 0 <= vec_idx < self.n_vec:
            raise IndexError(vec_idx)
        if not self._decoded:
            self._decode()
        out = []
        for c, dt in enumerate(self.dtypes):
            o = self.rg.offsets[c] + vec_idx * 1024 * dt.itemsize
            t = self.d_out[o: o + 1024 * dt.itemsize]
            out.append(t.view(getattr(torch, _TORCH_DT[dt.name])))
        return Chunk(out)

    def reset(self):
        self._decoded = False

    def column(self, c: int) -> np.ndarray:
        """Host copy of a whole decoded column."""
        if not self._decoded:
            self._decode()
        dt = self.dtypes[c]
        n = self.n_vec * 1024
        o = self.rg.offsets[c]
        return self.d_out[o: o + n * dt.itemsize].cpu().numpy().view(dt)
