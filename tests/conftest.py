"""Shared fixtures: the oracle, the compiled reference (when built) and the
GPU library, all driven through the em2d C API."""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_em2d.so")
GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")


def _load_oracle() -> C.CDLL:
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    lib = C.CDLL(ORACLE_LIB)
    lib.orc_track_ids.argtypes = [C.c_void_p]
    lib.orc_ids.restype = C.POINTER(C.c_int64)
    lib.orc_ids.argtypes = [C.c_void_p]
    return lib


@pytest.fixture(scope="session")
def oracle():
    return _load_oracle()


@pytest.fixture(scope="session")
def reference():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference build oracle/_ref is not present (needs the reference sources)")
    return C.CDLL(REF_LIB)


@pytest.fixture(scope="session")
def gpu():
    from zpic_amd import _lib

    return _lib.load()


def golden_cases() -> list[str]:
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load_golden(name: str):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    cfg = json.loads(str(z["cfg"]))
    return cfg, [int(s) for s in z["steps"]], z


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (n if n > 0 else 1.0))


def field_err(sa: dict, sb: dict, f: str) -> float:
    """Relative L2 difference of grid f between two states.  For B the
    reference norm is max(|B|, |E|): early in a run B is a small, noisy
    curl of E and its own norm would turn FP32 rounding of E into large
    relative numbers."""
    if f != "B":
        return rel_l2(sa[f], sb[f])
    b = np.asarray(sb["B"], dtype=np.float64)
    e = np.asarray(sb["E"], dtype=np.float64)
    d = np.linalg.norm((np.asarray(sa["B"], dtype=np.float64) - b).ravel())
    return float(d / max(np.linalg.norm(b.ravel()), np.linalg.norm(e.ravel()), 1e-300))


def oracle_ids(lib, run, k: int) -> np.ndarray:
    n = run.species[k].np
    return np.ctypeslib.as_array(lib.orc_ids(C.byref(run.species[k])), shape=(n,)).copy()


def gpu_ids(lib, run, k: int) -> np.ndarray:
    n = run.species[k].np
    ids = np.zeros(n, np.int64)
    lib.spec_get_ids(C.byref(run.species[k]), ids.ctypes.data_as(C.POINTER(C.c_int64)))
    return ids
