"""The drop-in boundary, checked on the CPU: the library loads, exports every
function the headers declare, its struct layouts match the ctypes mirror
and the reference layout plus one pointer, the host-side initialisation is
bit-identical to the reference's, and fatal input errors exit like the
reference does.  No GPU is touched here."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from conftest import ROOT, golden_cases, load_golden
from em2d_driver import Em2dRun, weibel
from zpic_amd import abi as A

INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "zpic_amd", "libzpic_em2d.so")

# declared in simulation.h but supplied by the user's deck (em2d/simulation.h:40,52)
USER_CALLBACKS = {"sim_init", "sim_report"}


def declared_functions() -> set[str]:
    names = set()
    for h in os.listdir(INCLUDE):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "#")):
                names.add(m.group(1))
    return names - USER_CALLBACKS


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libzpic_em2d.so not built")
    return C.CDLL(LIB)


def test_library_exports_every_declared_function(lib):
    names = declared_functions()
    assert {"sim_iter", "spec_advance", "current_update", "emf_advance", "current_zero"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert missing == []


def test_headers_compile_as_c99_and_layouts_match_ctypes(tmp_path):
    """sizeof/offsetof from the C headers equal the ctypes mirror, and equal
    the reference layout (em2d/particles.h:85-132 etc.) plus the trailing
    device pointer."""
    fields = {"t_species": ["part", "np", "energy", "density", "n_sort", "dev"],
              "t_emf": ["E_buf", "nrow", "ext_fld", "n_move", "dev"],
              "t_current": ["J_buf", "smooth", "moving_window", "dev"],
              "t_simulation": ["species", "emf", "current", "moving_window"],
              "t_part": ["ix", "uz"], "t_density": ["custom_x_total_q"]}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "simulation.h"', '#include "zpic_gpu.h"',
           "int main(void){"]
    for t, fs in fields.items():
        src.append(f'printf("{t} %zu\\n", sizeof({t}));')
        for f in fs:
            src.append(f'printf("{t}.{f} %zu\\n", offsetof({t}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{INCLUDE}", str(c), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    ours, ref = A.make_abi(True), A.make_abi(False)
    cls = {"t_species": ours.t_species, "t_emf": ours.t_emf, "t_current": ours.t_current,
           "t_simulation": ours.t_simulation, "t_part": A.t_part, "t_density": A.t_density}
    for t, fs in fields.items():
        assert int(got[t]) == C.sizeof(cls[t]), t
        for f in fs:
            assert int(got[f"{t}.{f}"]) == getattr(cls[t], f).offset, f"{t}.{f}"
    for t in ("t_species", "t_emf", "t_current"):
        assert getattr(ours, t).dev.offset >= C.sizeof(getattr(ref, t)) - 8


def test_em2d_program_builds_against_the_library(tmp_path):
    """A C program written only against the em2d API (the CPU timing driver,
    also built against the reference sources) links against the library."""
    exe = tmp_path / "drv"
    subprocess.run(["gcc", "-std=c99", f"-I{INCLUDE}", os.path.join(ROOT, "oracle", "bench_driver.c"),
                    f"-L{os.path.dirname(LIB)}", "-lzpic_em2d", f"-Wl,-rpath,{os.path.dirname(LIB)}", "-lm",
                    "-o", str(exe)], check=True)
    assert exe.exists()


@pytest.mark.parametrize("case", golden_cases())
def test_host_initial_state_matches_reference(lib, case):
    """spec_new injection, laser, initial and external fields run on the
    host (particles.c:85-609, emf.c:158-350, 764-980): the state before the
    first step is bit-identical to the reference's."""
    cfg, steps, z = load_golden(case)
    if 0 not in steps:
        pytest.skip("no initial state stored for this case")
    run = Em2dRun(lib, cfg)
    try:
        nx = tuple(cfg["nx"])
        assert A.grid_view(run.sim.emf.E_buf, nx).tobytes() == z["0/E"].tobytes()
        assert A.grid_view(run.sim.emf.B_buf, nx).tobytes() == z["0/B"].tobytes()
        for k in range(len(cfg["species"])):
            assert A.particles_view(run.species[k]).tobytes() == z[f"0/part{k}"].tobytes()
    finally:
        run.close()


def _child(code: str) -> subprocess.CompletedProcess:
    prog = textwrap.dedent(f"""
        import sys, ctypes as C
        sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
        from em2d_driver import Em2dRun, weibel
        from zpic_amd import abi as A
        lib = C.CDLL({LIB!r})
    """) + textwrap.dedent(code)
    return subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("what,code,message", [
    ("courant", "Em2dRun(lib, weibel(8, 8, 1, dt=0.08))", "courant condition violation"),
    ("smooth", "Em2dRun(lib, dict(weibel(8, 8, 1), smooth={'xtype': 'binomial', 'xlevel': 0}))",
     "Invalid smooth level along x"),
    ("laser", "cfg = weibel(8, 8, 1); cfg['laser'] = {'type': 'plane', 'fwhm': -1.0}; Em2dRun(lib, cfg)",
     "Invalid laser FWHM"),
    ("init_after_start", """
        r = Em2dRun(lib, weibel(8, 8, 1))
        r.sim.emf.iter = 1
        f = A.t_emf_init_fld()
        lib.emf_init_fld(C.byref(r.sim.emf), C.byref(f))""", "should only be called at initialization"),
    ("bad_particle", """
        r = Em2dRun(lib, weibel(8, 8, 1))
        r.species[0].part[3].ix = 99
        lib.spec_advance(C.byref(r.species[0]), C.byref(r.sim.emf), C.byref(r.sim.current))""", "outside the 8 x 8 grid"),
    ("batch_moving_window", """
        r = Em2dRun(lib, dict(weibel(8, 8, 1), moving_window=True))
        lib.zpic_batch_new.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        lib.zpic_batch_new((C.c_void_p * 1)(C.addressof(r.sim)), 1)""", "uses a moving window"),
    ("batch_ext_fld", """
        r = Em2dRun(lib, dict(weibel(8, 8, 1), ext_fld={"B_0": [0, 0, 0.1]}))
        lib.zpic_batch_new.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        lib.zpic_batch_new((C.c_void_p * 1)(C.addressof(r.sim)), 1)""", "external fields"),
    ("batch_twice", """
        r = Em2dRun(lib, weibel(8, 8, 1))
        lib.zpic_batch_new.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        lib.zpic_batch_new((C.c_void_p * 2)(C.addressof(r.sim), C.addressof(r.sim)), 2)""", "appears twice"),
])
def test_fatal_errors_exit_like_the_reference(lib, what, code, message):
    """Fatal input errors print a message and exit(-1) (simulation.c:106-110,
    136-144; emf.c:246-270, 924-927); particles outside the grid are refused
    before reaching the device."""
    out = _child(code)
    assert out.returncode == 255, (what, out.returncode, out.stderr[-500:])
    assert message in out.stderr + out.stdout


def test_python_module_imports_and_mirrors_reference_api():
    from zpic_amd import em2d

    for name in ("Density", "Species", "ExternalField", "InitialField", "EMF", "Laser", "Current", "Smooth",
                 "Simulation"):
        assert hasattr(em2d, name)
    s = em2d.Smooth(xtype="binomial", ytype="compensated", xlevel=2, ylevel=1)
    assert (s.xtype, s.ytype, s.xlevel) == ("binomial", "compensated", 2)
    assert s.ylevel == 2  # the reference getter returns xlevel (em2d.pyx:1720)
    d = em2d.Density(type="step", start=3.0)
    assert (d.type, d.start, d.n) == ("step", 3.0, 1.0)
    sp = em2d.Species("e", -1.0, [2, 2], n_sort=8)
    assert sp.n_sort == 8
