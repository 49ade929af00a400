"""Reader.get_chunk over a row-group directory: every vector of every column equals the written values."""
import numpy as np
import pytest

from fastlanes_amd import workloads as wl
from fastlanes_amd.reader import Connection

pytestmark = pytest.mark.gpu


def test_reader_get_chunk(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    n = 40 * 1024
    cols = [wl.rowgroup_column("bitpack", 8, n, 1), wl.rowgroup_column("delta", 64, n, 2),
            wl.rowgroup_column("rle", 32, n, 3), wl.rowgroup_column("dict", 16, n, 4),
            wl.geometric_runs(n, 6.0, seed=5, value_bits=10).astype(np.int64), np.linspace(0, 1, n),
            wl.dict_runs(n, np.uint16, seed=6)]
    exprs = [1, 11, 6, 17, 7, 0, 27]
    Connection.to_fls(str(tmp_path), cols, exprs)
    rd = Connection().read_fls(str(tmp_path))
    assert rd.n_vec == 40
    for v in (0, 17, 39):
        ch = rd.get_chunk(v)
        assert len(ch) == len(cols)
        for c, x in enumerate(cols):
            np.testing.assert_array_equal(ch[c].cpu().numpy(), x[v * 1024:(v + 1) * 1024], err_msg=f"col {c} vec {v}")
    for c, x in enumerate(cols):
        np.testing.assert_array_equal(rd.column(c), x)
    with pytest.raises(IndexError):
        rd.get_chunk(40)
