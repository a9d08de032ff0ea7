"""Write the golden vectors of the em2d time step from the reference itself.

Runs the reference em2d, compiled from its own sources by oracle/Makefile
(`make -C oracle ref`, strict IEEE flags) into oracle/_ref/libref_em2d.so,
through tests/em2d_driver.py, and stores the initial state and the states
after selected steps as compressed .npz files next to this script.  Only
these data files are committed; the reference cannot be loaded where the
tests run on the GPU.

Every case injects an even number of particles, so each run draws an even
number of normal deviates and leaves no cached deviate behind in the
process-global generator (em2d/random.c:69-70) for the next case.

Each file holds:  cfg (JSON of the run), steps, and for every dumped step
s: s/E, s/B, s/J (guard-inclusive float3 grids), s/part<k> (t_part
records), s/emf_energy, s/kin_energy, s/n_move.

usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from em2d_driver import Em2dRun, lwfa, weibel  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_em2d.so")


def cases() -> dict:
    """name -> (config, dumped steps, dump particles?)"""
    c = {}
    # reference Weibel deck (electrons + positrons), 2x2 ppc
    c["weibel_32"] = (weibel(32, 32, 2), [0, 1, 10, 17], True)
    c["weibel_32_long"] = (weibel(32, 32, 2), [100], False)
    # two electron populations (the benchmark configurations), 4x4 ppc
    two_e = weibel(24, 16, 4, n_sort=20)
    two_e["species"][1]["m_q"] = -1.0
    two_e["species"][1]["name"] = "electrons2"
    c["weibel2e_24x16"] = (two_e, [0, 1, 21], True)
    # smoothing: binomial x/y; compensated x level 4; y passes counted with xlevel
    c["smooth_binomial"] = (weibel(32, 24, 2, smooth={"xtype": "binomial", "ytype": "binomial",
                                                      "xlevel": 1, "ylevel": 1}), [1, 10], True)
    c["smooth_compensated"] = (weibel(32, 32, 2, smooth={"xtype": "compensated", "xlevel": 4}), [1, 10], False)
    c["smooth_ylevel_quirk"] = (weibel(32, 32, 2, smooth={"xtype": "binomial", "ytype": "compensated",
                                                          "xlevel": 2, "ylevel": 1}), [1, 10], False)
    # moving window, step density, Gaussian laser, compensated smoothing
    c["lwfa_200x32"] = (lwfa(200, 32, start=6.0), [0, 1, 60], True)
    # uniform initial fields: gyration and E x B drift of a cold sparse plasma
    cold = weibel(16, 16, 1, uz=0.3, uth=0.0)
    cold["init_fld"] = {"E_0": [0.0, 0.05, 0.0], "B_0": [0.0, 0.0, 1.0]}
    c["uniform_fields"] = (cold, [0, 1, 10, 50], True)
    # uniform external fields
    ext = weibel(16, 16, 2)
    ext["ext_fld"] = {"E_0": [0.01, 0.0, 0.0], "B_0": [0.0, 0.0, 0.5]}
    c["external_fields"] = (ext, [1, 10], True)
    return c


def dump(run: Em2dRun, step: int, out: dict) -> None:
    st = run.state()
    for k, v in st.items():
        out[f"{step}/{k}"] = np.asarray(v)


def main() -> int:
    if not os.path.exists(REF_LIB):
        print(f"{REF_LIB} missing: run `make -C oracle ref` first", file=sys.stderr)
        return 1
    lib = C.CDLL(REF_LIB)
    for name, (cfg, steps, with_parts) in cases().items():
        run = Em2dRun(lib, cfg, with_dev=False)
        out = {"cfg": np.array(json.dumps(cfg)), "steps": np.array(steps, dtype=np.int64)}
        for s in steps:
            run.step(s - run.n)
            dump(run, s, out)
            if not with_parts:
                for k in [k for k in out if k.startswith(f"{s}/part")]:
                    out[k] = np.zeros(0, dtype=out[k].dtype)
        run.close()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: steps {steps} -> {os.path.getsize(path) / 1024:.0f} KiB")
    return 0


if __name__ == "__main__":
    sys.exit(main())
