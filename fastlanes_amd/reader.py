"""Connection / Reader mirror for the decode path: a row-group written to a directory and read back through one
device-side decode.

Follows the reference's file path (`Connection::to_fls` / `Connection::read_fls`, src/connection.cpp:102-128;
`Reader(dir, Connection&)`, `Reader::get_chunk(vec_idx)`, src/reader/reader.cpp:19-61; footer fields of
include/fls/footer/rowgroup_footer.hpp:23-75): a directory holds `data.fls` (the encoded columns back to back) and
`footer.json` (per column: data type, byte offset and size in data.fls, expression id; plus n_vec and
rowgroup_size). Differences: the footer stores this library's expression id instead of an RPN, and `get_chunk`
returns device tensors of the whole row-group decoded by one persistent launch (replayed from a HIP graph) rather
than decoding vector by vector on the host.
"""
from __future__ import annotations

import json
import os

import numpy as np

import fastlanes_amd as fa

FASTLANES_FILE_NAME = "data.fls"
FOOTER_FILE_NAME = "footer.json"
_TORCH_DT = {"int8": "int8", "int16": "int16", "int32": "int32", "int64": "int64", "uint8": "uint8",
             "uint16": "uint16", "uint32": "uint32", "uint64": "uint64", "float64": "float64"}


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
                metas.append({"data_type": c.dtype.name, "offset": pos, "size": int(enc.size), "expression": int(eid)})
                f.write(enc.tobytes())
                pos += enc.size
        footer = {"n_vec": (n.pop() if n else 0) // 1024, "rowgroup_size": pos, "column_metadatas": metas}
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
        if data.size != self.footer["rowgroup_size"]:
            raise fa.FlsError(f"{dir_path}: data size {data.size} != footer rowgroup_size")
        self.device = device or torch.device("cuda:0")
        self.stream = stream
        metas = self.footer["column_metadatas"]
        self.dtypes = [np.dtype(m["data_type"]) for m in metas]
        cols = [data[m["offset"]: m["offset"] + m["size"]] for m in metas]
        self.rg = fa.RowGroup(cols, self.dtypes, [m["offset"] for m in metas])
        self.d_enc = torch.from_numpy(np.concatenate([data, np.zeros(16, np.uint8)])).to(self.device)
        self.d_out = torch.empty(self.rg.total_bytes + 256, dtype=torch.uint8, device=self.device)
        self._decoded = False
        self.n_vec = int(self.footer["n_vec"])

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
