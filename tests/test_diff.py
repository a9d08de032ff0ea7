"""The FP64 differentiable restatement (oracle/diff64.py) that checks the
GPU adjoint: pinned against the C oracle's time step (float32 runs agree),
and its autograd gradients against central finite differences."""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from conftest import ROOT, oracle_ids
from em2d_driver import Em2dRun, weibel

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import diff64  # noqa: E402

from zpic_amd import abi as A  # noqa: E402


def torch_species(parts, qs, m_qs, dtype, uz=None, uth=None, G=None):
    out = []
    for k, p in enumerate(parts):
        if G is None:
            u = torch.tensor(np.stack([p["ux"], p["uy"], p["uz"]], 1), dtype=dtype)
        else:
            sign = 1.0 if k == 0 else -1.0
            g = torch.tensor(G[k], dtype=dtype)
            u = uth * g + torch.stack([torch.zeros((), dtype=dtype), torch.zeros((), dtype=dtype), sign * uz]) \
                if torch.is_tensor(uz) else uth * g + torch.tensor([0.0, 0.0, sign * uz], dtype=dtype)
        out.append(diff64.Species(qs[k], m_qs[k], torch.tensor(p["ix"]), torch.tensor(p["iy"]),
                                  torch.tensor(p["x"], dtype=dtype), torch.tensor(p["y"], dtype=dtype), u))
    return out


def grid_of(run):
    nx, ny = run.cfg["nx"]
    E = torch.tensor(np.ctypeslib.as_array(C.cast(run.sim.emf.E_buf, C.POINTER(C.c_float)),
                                           shape=(ny + 3, nx + 3, 3)).copy())
    B = torch.tensor(np.ctypeslib.as_array(C.cast(run.sim.emf.B_buf, C.POINTER(C.c_float)),
                                           shape=(ny + 3, nx + 3, 3)).copy())
    return E, B


def test_float32_restatement_follows_the_oracle(oracle):
    """Eight steps of a 32x24 Weibel run: the torch restatement in float32
    and the C oracle keep identical cells and agree to FP32 rounding."""
    cfg = weibel(32, 24, 2, n_sort=0, uth=0.1, uz=0.6)
    run = Em2dRun(oracle, cfg)
    for k in range(2):
        oracle.orc_track_ids(C.byref(run.species[k]))
    try:
        nx, ny = cfg["nx"]
        grid = diff64.Grid(nx, ny, 0.1, 0.1, cfg["dt"], dtype=torch.float32)
        parts = [A.particles_view(run.species[k]).copy() for k in range(2)]
        sp = torch_species(parts, [run.species[k].q for k in range(2)], [run.species[k].m_q for k in range(2)],
                           torch.float32)
        E, B = grid_of(run)
        for n in range(8):
            E, B = diff64.step(grid, sp, E, B)
            run.step(1)
        st = run.state()
        for k in range(2):
            ids = oracle_ids(oracle, run, k)
            p = st[f"part{k}"][np.argsort(ids)]
            assert np.array_equal(p["ix"], sp[k].ix.numpy()) and np.array_equal(p["iy"], sp[k].iy.numpy())
            assert np.allclose(p["x"], sp[k].x.numpy(), atol=2e-6)
            assert np.allclose(p["uz"], sp[k].u[:, 2].numpy(), atol=2e-6)
        for name, T in (("E", E), ("B", B)):
            ref = st[name]  # guard-inclusive, as the torch grids
            got = T.numpy()
            scale = max(np.abs(st["E"]).max(), np.abs(st["B"]).max())
            assert np.abs(got - ref).max() <= 1e-5 * scale, name
    finally:
        run.close()


def test_fp64_gradients_match_central_differences(oracle):
    """Autograd through the FP64 restatement against central differences
    for the drift, the thermal spread and initial B_z nodes.  The step is
    piecewise smooth (kinks where particles cross interpolation nodes and
    cell edges), so the difference step is small: h = 1e-6 in FP64."""
    cfg = weibel(16, 16, 2, n_sort=0)
    run = Em2dRun(oracle, dict(cfg, species=[dict(s, ufl=[0, 0, 0], uth=[1, 1, 1]) for s in cfg["species"]]))
    try:
        parts = [A.particles_view(run.species[k]).copy() for k in range(2)]
        G = [np.stack([p["ux"], p["uy"], p["uz"]], 1).astype(np.float64) for p in parts]
        qs = [run.species[k].q for k in range(2)]
    finally:
        run.close()
    grid = diff64.Grid(16, 16, 0.1, 0.1, 0.07)
    rng = np.random.default_rng(3)
    bz0 = torch.tensor(1e-2 * rng.standard_normal((19, 19)))

    def loss(uz, uth, bz):
        sp = torch_species(parts, qs, [-1.0, -1.0], torch.float64, uz=uz, uth=uth, G=G)
        B = grid.zeros()
        B = B + torch.nn.functional.pad(bz[..., None], (2, 0))
        E = grid.zeros()
        return diff64.run_loss(grid, sp, E, B, 10, 4, 10)[0]

    uz = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    uth = torch.tensor(0.08, dtype=torch.float64, requires_grad=True)
    bz = bz0.clone().requires_grad_(True)
    L = loss(uz, uth, bz)
    L.backward()
    h = 1e-6
    with torch.no_grad():
        fd_uz = (loss(uz + h, uth, bz) - loss(uz - h, uth, bz)) / (2 * h)
        fd_uth = (loss(uz, uth + h, bz) - loss(uz, uth - h, bz)) / (2 * h)
        assert abs(float(uz.grad) - float(fd_uz)) <= 1e-5 * abs(float(fd_uz)) + 1e-12
        assert abs(float(uth.grad) - float(fd_uth)) <= 1e-5 * abs(float(fd_uth)) + 1e-12
        for (j, i) in ((5, 7), (0, 3), (17, 18)):
            e = torch.zeros_like(bz)
            e[j, i] = h
            fd = (loss(uz, uth, bz + e) - loss(uz, uth, bz - e)) / (2 * h)
            assert abs(float(bz.grad[j, i]) - float(fd)) <= 1e-5 * abs(float(fd)) + 1e-12, (j, i)
