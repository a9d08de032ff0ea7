"""Compare the oracle with the compiled reference on one case, in a fresh
process (the reference's generator keeps a cached deviate between runs,
em2d/random.c:69-70, so every comparison starts from a fresh library).

usage: python tests/ref_worker.py <case>   -> prints a JSON result
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from em2d_driver import Em2dRun, lwfa, weibel  # noqa: E402


def cases() -> dict:
    slab = weibel(40, 20, 3, uth=0.2)
    slab["species"][0]["density"] = {"type": "slab", "start": 1.0, "end": 3.0}
    custom = weibel(24, 20, 2, uth=0.05)
    custom["species"][1]["density"] = {"type": "custom",
                                       "custom_x": lambda x: 1.0 + 0.5 * math.sin(x),
                                       "custom_y": lambda y: 1.0 + 0.25 * math.cos(2 * y)}
    fields = weibel(20, 20, 2)
    fields["init_fld"] = {"E_custom": lambda i, dx, j, dy: (0.01 * math.sin(i * dx), 0.0, 0.02 * math.cos(j * dy))}
    fields["ext_fld"] = {"B_custom": lambda i, dx, j, dy: (0.0, 0.0, 0.3 + 0.1 * math.sin(i * dx + j * dy))}
    plane = lwfa(120, 16, start=3.0, smooth_level=2)
    plane["laser"]["type"] = "plane"
    plane["laser"]["polarization"] = 0.3
    odd = weibel(27, 13, 3, uth=0.15, smooth={"xtype": "compensated", "ytype": "binomial", "xlevel": 1,
                                             "ylevel": 2})
    return {"slab": (slab, 25), "custom_density": (custom, 20), "custom_fields": (fields, 20),
            "plane_laser": (plane, 40), "odd_grid": (odd, 30)}


def report_case():
    """A deck with non-trivial initial fields (Gaussian laser), a STEP
    plasma, moving window and external-free particle fields."""
    cfg = lwfa(120, 16, start=3.0, smooth_level=2)
    cfg["species"][0]["ppc"] = [2, 2]
    return cfg


PHA_NX = (32, 16)
PHA_RANGE = ((0.0, 6.0), (-0.05, 0.05))


def write_reports(lib, run, outdir: str) -> None:
    """Every report of the em2d API at the run's current step, into outdir."""
    cwd = os.getcwd()
    os.makedirs(outdir, exist_ok=True)
    os.chdir(outdir)
    try:
        for field in range(4):  # EFLD, BFLD, EPART, BPART
            for fc in range(3):
                lib.emf_report(C.byref(run.sim.emf), C.c_byte(field), fc)
        for jc in range(3):
            lib.current_report(C.byref(run.sim.current), jc)
        nx = (C.c_int * 2)(*PHA_NX)
        rng = ((C.c_float * 2) * 2)((C.c_float * 2)(*PHA_RANGE[0]), (C.c_float * 2)(*PHA_RANGE[1]))
        sp = C.byref(run.species[0])
        lib.spec_report(sp, 0x3000, None, None)                     # PARTICLES
        lib.spec_report(sp, 0x2000 + 1 + 4 * 16, nx, rng)           # PHASESPACE(X1, U1)
        lib.spec_report(sp, 0x2000 + 2 + 6 * 16, nx, rng)           # PHASESPACE(X2, U3)
        if os.environ.get("ZPIC_REPORT_CHARGE"):
            lib.spec_report(sp, 0x1000, None, None)                 # CHARGE
    finally:
        os.chdir(cwd)


def main(name: str) -> int:
    cfg, steps = cases()[name]
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_em2d.so"))
    orc = C.CDLL(os.path.join(ROOT, "oracle", "_build", "liboracle.so"))
    a = Em2dRun(ref, cfg, with_dev=False)
    b = Em2dRun(orc, cfg)
    for _ in range(steps):
        a.step()
        b.step()
    sa, sb = a.state(), b.state()
    bad = [k for k in sa if np.asarray(sa[k]).tobytes() != np.asarray(sb[k]).tobytes()]
    print(json.dumps({"case": name, "mismatch": bad, "np": [int(len(sa[k])) for k in sa if k.startswith("part")]}))
    a.close()
    b.close()
    return 0


def main_report(outdir: str, steps: int = 0) -> int:
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_em2d.so"))
    run = Em2dRun(ref, report_case(), with_dev=False)
    run.step(steps)
    write_reports(ref, run, outdir)
    run.close()
    print(json.dumps({"report": outdir}))
    return 0


if __name__ == "__main__":
    if sys.argv[1] == "report":
        sys.exit(main_report(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 0))
    sys.exit(main(sys.argv[1]))
