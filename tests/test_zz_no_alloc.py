"""spec_deposit_charge_noalloc (no counterpart in the reference): the
allocation-free variant of spec_deposit_charge gives the same density as
that entry point, call after call (the second call on a smaller grid, so a
larger call's leftovers in the shared scratch would show), adds to the
caller's array like it, and refuses grids beyond its scratch and species
without device state, without touching the caller's array."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2
from em2d_driver import Em2dRun, weibel

pytestmark = pytest.mark.gpu

FP = C.POINTER(C.c_float)


def densities(gpu, run, k, start):
    nx, ny = run.cfg["nx"]
    a = np.full((ny + 1) * (nx + 1), start, np.float32)
    b = a.copy()
    gpu.spec_deposit_charge(C.byref(run.species[k]), a.ctypes.data_as(FP))
    rc = gpu.spec_deposit_charge_noalloc(C.byref(run.species[k]), b.ctypes.data_as(FP))
    return rc, a, b


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_charge_matches_spec_deposit_charge(gpu, order):
    cfgs = [weibel(96, 80, 3, uth=0.2, n_sort=7, m_q2=-1.0), weibel(24, 16, 2, uth=0.1)]
    for n, cfg in enumerate(cfgs):
        run = Em2dRun(gpu, cfg, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))
        try:
            run.step(9)
            for k in range(2):
                for start in (0.0, 0.25):
                    rc, a, b = densities(gpu, run, k, start)
                    print(f"noalloc: call {n} species {k} start {start}: rc {rc}, rel {rel_l2(b, a):.2e}",
                          flush=True)
                    assert rc == 0
                    assert rel_l2(b, a) <= 1e-6
                    assert np.abs(b - a).max() <= 1e-6 * np.abs(a).max()
        finally:
            run.close()


def test_noalloc_charge_refuses_without_touching_the_array(gpu):
    # a species that was never on the device: refused before any GPU work
    cfg = weibel(16, 16, 2)
    run = Em2dRun(gpu, cfg, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))
    try:
        out = np.full(17 * 17, 3.0, np.float32)
        assert gpu.spec_deposit_charge_noalloc(C.byref(run.species[0]), out.ctypes.data_as(FP)) == -1
        assert np.all(out == 3.0)
    finally:
        run.close()
    # a grid of more charge nodes than the scratch holds: refused on size
    big = {"nx": [2900, 2900], "box": [290.0, 290.0], "dt": 0.07,
           "species": [{"name": "sparse", "m_q": -1.0, "ppc": [1, 1],
                        "density": {"type": "step", "start": 1000.0, "n": 1.0}}]}
    run = Em2dRun(gpu, big, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))
    try:
        out = np.zeros(2901 * 2901, np.float32)
        assert gpu.spec_deposit_charge_noalloc(C.byref(run.species[0]), out.ctypes.data_as(FP)) == -1
        assert not np.any(out)
    finally:
        run.close()
