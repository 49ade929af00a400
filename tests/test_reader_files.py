"""Row-group files (data.fls + footer.json) written by Connection.to_fls: layout checks and oracle decode of every
column slice (CPU)."""
import json
import os

import numpy as np

from fastlanes_amd import workloads as wl
from fastlanes_amd.reader import FASTLANES_FILE_NAME, FOOTER_FILE_NAME, Connection


def test_to_fls_layout_and_oracle_decode(tmp_path, oracle):
    n = 8 * 1024
    cols = [wl.rowgroup_column("ffor", 32, n, 1), wl.timestamps(n), wl.rowgroup_column("rle", 16, n, 2),
            wl.rowgroup_column("dict", 8, n, 3), np.arange(n, dtype=np.float64), wl.dict_runs(n, np.int32, seed=4)]
    exprs = [1, 11, 6, 17, 0, 27]
    footer = Connection.to_fls(str(tmp_path), cols, exprs)
    assert footer["n_vec"] == 8
    on_disk = json.load(open(os.path.join(tmp_path, FOOTER_FILE_NAME)))
    assert on_disk == footer
    data = np.fromfile(os.path.join(tmp_path, FASTLANES_FILE_NAME), dtype=np.uint8)
    assert data.size == footer["rowgroup_size"]
    for x, m in zip(cols, footer["column_metadatas"]):
        assert m["offset"] % 256 == 0 and m["data_type"] == x.dtype.name
        got = oracle.decode_column(data[m["offset"]: m["offset"] + m["size"]], x.dtype, x.size)
        np.testing.assert_array_equal(got, x)
