"""Row-group files (data.fls + footer.json) written by Connection.to_fls: the footer in the reference's keys
(src/json/fls_json.cpp:30-61), layout checks and oracle decode of every column slice (CPU)."""
import json
import os

import numpy as np

from fastlanes_amd import workloads as wl
from fastlanes_amd.reader import FASTLANES_FILE_NAME, FOOTER_FILE_NAME, Connection, column_dtypes, column_sizes


def test_to_fls_layout_and_oracle_decode(tmp_path, oracle):
    n = 8 * 1024
    cols = [wl.rowgroup_column("ffor", 32, n, 1), wl.timestamps(n), wl.rowgroup_column("rle", 16, n, 2),
            wl.rowgroup_column("dict", 8, n, 3), np.arange(n, dtype=np.float64), wl.dict_runs(n, np.int32, seed=4)]
    exprs = [1, 11, 6, 17, 0, 27]
    footer = Connection.to_fls(str(tmp_path), cols, exprs)
    assert set(footer) == {"1: N VEC", "2: Rowgroup size", "3: Column Metadatas"}
    assert footer["1: N VEC"] == 8
    on_disk = json.load(open(os.path.join(tmp_path, FOOTER_FILE_NAME)))
    assert on_disk == footer
    data = np.fromfile(os.path.join(tmp_path, FASTLANES_FILE_NAME), dtype=np.uint8)
    assert data.size == footer["2: Rowgroup size"]
    names = {1: "DOUBLE", 4: "INT32", 5: "INT64", 6: "UINT8", 7: "UINT16", 8: "UINT32"}
    exp = {0: "UNCOMPRESSED", 1: "FFOR", 6: "RLE", 11: "DELTA", 17: "DICT", 27: "DICT_RLE"}
    for c, (x, eid, m) in enumerate(zip(cols, exprs, footer["3: Column Metadatas"])):
        assert set(m) >= {"logical_type", "logical_type_str", "Encoding RPN", "col_ofs", "is_determined"}
        assert m["col_ofs"] % 256 == 0 and m["logical_type_str"] == names[m["logical_type"]]
        assert m["Encoding RPN"] == {"operators": [exp[eid], m["logical_type_str"]], "operands": [c]}
        got = oracle.decode_column(data[m["col_ofs"]: m["col_ofs"] + m["size"]], x.dtype, x.size)
        np.testing.assert_array_equal(got, x)
    assert column_dtypes(footer) == [x.dtype for x in cols]
    # without "size", a column runs to the next column (or the end): the oracle decodes the padded slices too
    for m in footer["3: Column Metadatas"]:
        del m["size"]
    for x, m, sz in zip(cols, footer["3: Column Metadatas"], column_sizes(footer)):
        np.testing.assert_array_equal(oracle.decode_column(data[m["col_ofs"]: m["col_ofs"] + sz], x.dtype, x.size), x)
