"""Host-mirror views of the em2d module (no GPU): reads leave the mirror
alone, every numpy-level write refreshes it first and marks it edited."""
from __future__ import annotations

import numpy as np

from zpic_amd.em2d import _watched


def make(shape=(4, 5), dtype=np.float32):
    log = []
    base = np.zeros(shape, dtype)
    return base, _watched(base, lambda: log.append("sync"), lambda: log.append("edit")), log


def test_reads_do_not_mark_edits():
    base, v, log = make()
    _ = v + 1, v.sum(), v.max(), np.abs(v), v[1:3], v.copy(), float(v[0, 0]), v.T, np.mean(v)
    assert log == []
    assert type(v + 1) is np.ndarray  # results of arithmetic are plain arrays


def test_writes_refresh_then_mark():
    base, v, log = make()
    v[1, 2] = 3
    assert log == ["sync", "edit"] and base[1, 2] == 3
    log.clear()
    v += 2
    assert log == ["sync", "edit"] and base[0, 0] == 2
    log.clear()
    np.multiply(v, 2, out=v)
    assert log == ["sync", "edit"] and base[0, 0] == 4
    log.clear()
    v.fill(7)
    np.copyto(v, 8)
    np.put(v, [0], [9])
    np.add.at(v, (0, 1), 1)
    assert log == ["sync", "edit"] * 4
    assert base[0, 0] == 9 and base[0, 1] == 9


def test_slices_and_fields_report_to_the_same_mirror():
    log = []
    base = np.zeros(6, dtype=[("x", "f4"), ("ix", "i4")])
    v = _watched(base, lambda: log.append("sync"), lambda: log.append("edit"))
    v["x"][2:4] = 1.5
    assert base["x"][2] == 1.5 and log[:2] == ["sync", "edit"]
    log.clear()
    sub = v[3:]
    sub["ix"] += 1
    assert base["ix"][3] == 1 and base["ix"][0] == 0 and "edit" in log
