"""Staged re-sorts against full re-sorts.

With zpic_set_staged_sort on, the push of a step that ends in a tile sort
writes the particles that stayed in their tile to the front of the tile's
range of the other buffer set and the others to a staging area; a second
pass moves only those to their new tiles.  The reference sorts by cell on
the host (em2d/particles.c:942-1007, every n_sort steps, :1262-1264); the
order of the particles is not part of any result, and the current is an
exact fixed-point sum, so a run with staged re-sorts must give the same E,
B and J bit for bit and the same particles (matched by id) as a run whose
every sort reads and writes all particles.

Mode 2 leaves one free slot per tile range, so ranges fill up: a staged
re-sort that cannot put every particle in its own tile's range (no room in
the staging area or in the new range) is followed by a full re-sort, and
the results are again those of full sorts bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from test_gpu_configs import _id_sorted
from em2d_driver import weibel
from test_gpu_parity import gpu_run

pytestmark = pytest.mark.gpu


def _run(gpu, cfg, mode, steps, open_y=False):
    gpu.zpic_set_staged_sort(mode)
    try:
        n0 = gpu.zpic_staged_sorts()
        r = gpu_run(gpu, cfg, ids=True)
        try:
            if open_y:
                gpu.zpic_sim_set_open_y(C.byref(r.sim))
            r.step(steps)
            staged = gpu.zpic_staged_sorts() - n0
            st = r.state()
            parts = [_id_sorted(gpu, r, k) for k in range(len(cfg["species"]))]
            return st, parts, staged
        finally:
            r.close()
    finally:
        gpu.zpic_set_staged_sort(0)  # the default


@pytest.mark.parametrize("mode", [1, 2])
def test_staged_sorts_match_full_sorts(gpu, mode):
    # hot electrons (u_th 0.2) and a sort every 3 steps: many particles
    # change tile between sorts; 48 steps, no host reads in between
    cfg = weibel(64, 48, 3, uth=0.2, n_sort=3, m_q2=-1.0)
    steps = 48
    full, pf, nf = _run(gpu, cfg, 0, steps)
    st, ps, ns = _run(gpu, cfg, mode, steps)
    print(f"mode {mode}: staged re-sorts {ns} (full-sort run: {nf})", flush=True)
    assert nf == 0
    # 16 scheduled sorts per species: the first is a full one, which makes
    # the ranges' free slots (the upload packs them), the others staged (in
    # mode 2 most of them fill ranges and are followed by a full re-sort)
    assert ns >= 2 * (steps // 3 - 1)
    # after either kind of sort every particle is in the range of its own
    # tile, so every tile deposits the same particles: J, E and B bit for bit
    for f in ("E", "B", "J"):
        assert st[f].tobytes() == full[f].tobytes(), f
    for k in range(2):
        assert np.array_equal(ps[k][1], pf[k][1])
        assert ps[k][0].tobytes() == pf[k][0].tobytes(), k


def test_staged_sorts_with_open_y_removals(gpu):
    """Open y boundaries remove the particles that leave through y: the
    sorting push drops them (they take no slot), and the staged runs match
    full re-sorts bit for bit, with the same particles left."""
    cfg = weibel(48, 40, 3, uth=0.3, n_sort=3, m_q2=-1.0)
    steps = 30
    full, pf, _ = _run(gpu, cfg, 0, steps, open_y=True)
    st, ps, ns = _run(gpu, cfg, 1, steps, open_y=True)
    print(f"open y: staged re-sorts {ns}", flush=True)
    assert ns >= 2 * (steps // 3 - 1)
    n0 = 48 * 40 * 9
    assert all(len(p[1]) < n0 for p in pf)  # particles did leave
    for f in ("E", "B", "J"):
        assert st[f].tobytes() == full[f].tobytes(), f
    for k in range(2):
        assert np.array_equal(ps[k][1], pf[k][1])
        assert ps[k][0].tobytes() == pf[k][0].tobytes(), k
