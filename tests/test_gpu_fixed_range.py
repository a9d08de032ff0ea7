"""Range of the push's fixed-point current windows on the GPU.

A tile's window sums its particles' contributions exactly while the tile
holds at most ZK_FX_TILE_MAX particles (all species of the launch; each
adds at most Q to a node).  The largest tile under the limit gives the
oracle's fields; a tile over it is counted by the push, and the library
stops the run at its next read of the species' counters, with a message
naming the limit.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import field_err
from em2d_driver import Em2dRun
from test_gpu_parity import gpu_run, oracle_run

pytestmark = pytest.mark.gpu

TILE_MAX = (1 << 18) - 16
HERE = os.path.dirname(os.path.abspath(__file__))


def one_tile(ppc):
    """a 16 x 16 grid: the default tiling makes it one tile"""
    return {"nx": [16, 16], "box": [1.6, 1.6], "dt": 0.07, "seed": [12345, 67890],
            "species": [{"name": "electrons", "m_q": -1.0, "ppc": list(ppc), "ufl": [0, 0, 0.6],
                         "uth": [0.1, 0.1, 0.1], "n_sort": 4}]}


def test_largest_exact_tile_matches_the_oracle(gpu, oracle):
    cfg = one_tile((31, 32))  # 253,952 particles in one tile
    assert 16 * 16 * 31 * 32 <= TILE_MAX
    g, o = gpu_run(gpu, cfg), oracle_run(oracle, cfg)
    try:
        print("one tile of 253952 particles: 5 steps", flush=True)
        g.step(5)
        o.step(5)
        sg, so = g.state(), o.state()
        for f in ("E", "B", "J"):
            assert field_err(sg, so, f) <= 1e-5, f
        assert np.allclose(sg["kin_energy"], so["kin_energy"], rtol=1e-6)
    finally:
        g.close()
        o.close()


SCRIPT = """
import ctypes as C, sys
sys.path.insert(0, {here!r})
from em2d_driver import Em2dRun
from test_gpu_fixed_range import one_tile
from zpic_amd import _lib
lib = _lib.load()
run = Em2dRun(lib, one_tile((32, 33)))
run.step(2)
lib.sim_sync_host(C.byref(run.sim))
print("no stop", flush=True)
"""


def test_tile_over_the_range_stops_the_run():
    assert 16 * 16 * 32 * 33 > TILE_MAX
    print("one tile of 270336 particles, in a child process", flush=True)
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([os.path.dirname(HERE), HERE, env.get("PYTHONPATH", "")])
    p = subprocess.run([sys.executable, "-c", SCRIPT.format(here=HERE)], capture_output=True, text=True,
                       timeout=120, env=env)
    assert p.returncode != 0, p.stdout + p.stderr
    assert "no stop" not in p.stdout
    assert "range of the fixed-point current windows" in p.stderr, p.stderr
