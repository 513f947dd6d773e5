"""Snapshot files (saveStats, StokesFlow2D_multixpu.jl:212-223) written by the library's MAT
level-5 writer and read back with scipy's independent MAT reader.  Host only."""
import numpy as np
import scipy.io

import sfhip


def test_save_stats_roundtrip(tmp_path):
    path = tmp_path / "stats.mat"
    sfhip.saveStats(path, (2, 3, 1), 17)
    raw = path.read_bytes()
    assert raw[:19] == b"MATLAB 5.0 MAT-file" and raw[126:128] == b"IM" and len(raw) % 8 == 0
    m = scipy.io.loadmat(str(path))
    assert m["dims"].dtype == np.int64 and m["dims"].shape == (3, 1)
    assert m["dims"].ravel().tolist() == [2, 3, 1]
    assert m["nt"].dtype == np.int64 and m["nt"].shape == (1, 1) and int(m["nt"][0, 0]) == 17


def test_save_stats_bad_path(tmp_path):
    try:
        sfhip.saveStats(tmp_path / "missing_dir" / "stats.mat", (1, 1, 1), 0)
    except sfhip.SFError as e:
        assert "cannot open" in str(e)
    else:
        raise AssertionError("expected an error for an unwritable path")
