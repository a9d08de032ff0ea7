"""Open (absorbing) y boundaries, zpic_sim_set_open_y.

em2d has no open boundary (its y boundary is always periodic), so nothing
in the reference fixes these results: they are pinned by mathematics.

* Fields: a vacuum pulse at normal incidence leaves through the boundary
  and the energy left behind is the reflected part.  For the first-order
  Mur condition E_g(n+1) = E_in(n) + S (E_in(n+1) - E_g(n)),
  S = (dt - dy)/(dt + dy), applied to a Yee plane wave
  E_j = z^n (a^j + R a^-j) (z = exp(i w dt), a = exp(i k dy), w from the
  discrete dispersion relation sin(w dt/2)/dt = sin(k dy/2)/dy), the
  reflection coefficient is
      R = (1 + S z - S/a - z/a) / (z a - 1 - S z + S a),
  about 5e-3 at 16 cells per wavelength.  The reflected energy must stay
  below twice the largest |R|^2 over the pulse's band (a periodic or a
  reflecting boundary leaves all of it).
* Particles: a cold, exactly neutral beam (electrons and positrons with the
  same positions and momenta: every current contribution cancels, fields
  stay exactly zero) moves ballistically; the particles removed at each
  step are exactly those whose cell left [0, ny), replayed in float32 with
  the push's own operations.
* Current: in a thermal plasma the discrete continuity equation holds at
  FP32 rounding on every node row that removals do not touch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import gpu_ids
from em2d_driver import Em2dRun, weibel
from zpic_amd import abi as A

pytestmark = pytest.mark.gpu


def mur_reflection(k, dt, d):
    S = np.float32((dt - d) / (dt + d))
    w = 2.0 / dt * np.arcsin(dt / d * np.sin(k * d / 2.0))
    z, a = np.exp(1j * w * dt), np.exp(1j * k * d)
    return np.abs((1 + S * z - S / a - z / a) / (z * a - 1 - S * z + S * a))


def pulse_cfg(nx, ny, d, dt, lam, width, direction):
    """Vacuum pulse of wavelength lam and Gaussian envelope `width` centred
    in y, uniform in x: (Ex, Bz) moving to -y, or (Ez, Bx) moving to +y
    (E and B at the same time level, each at its own Yee position)."""
    k0 = 2.0 * np.pi / lam
    y0 = ny * d / 2.0

    def f(y):
        return float(np.exp(-((y - y0) / width) ** 2) * np.sin(k0 * (y - y0)))

    if direction < 0:
        E = lambda i, dx, j, dy: (f(j * dy), 0.0, 0.0)
        B = lambda i, dx, j, dy: (0.0, 0.0, f((j + 0.5) * dy))
    else:
        E = lambda i, dx, j, dy: (0.0, 0.0, f(j * dy))
        B = lambda i, dx, j, dy: (f((j + 0.5) * dy), 0.0, 0.0)
    return {"nx": [nx, ny], "box": [nx * d, ny * d], "dt": dt, "species": [],
            "init_fld": {"E_custom": E, "B_custom": B}}


def field_energy(lib, run):
    e6 = (C.c_double * 6)()
    lib.emf_get_energy(C.byref(run.sim.emf), e6)
    return float(sum(e6[:]))


@pytest.mark.parametrize("direction", [-1, +1], ids=["down_Ex_Bz", "up_Ez_Bx"])
def test_vacuum_pulse_leaves_through_open_y(gpu, direction):
    nx, ny, d, dt = 16, 256, 0.1, 0.07
    lam, width = 16 * d, 2 * 16 * d
    cfg = pulse_cfg(nx, ny, d, dt, lam, width, direction)
    k0, sk = 2 * np.pi / lam, np.sqrt(2.0) / width
    band = np.linspace(k0 - 3 * sk, k0 + 3 * sk, 200)
    bound = 2.0 * float(np.max(mur_reflection(band, dt, d) ** 2))
    # the incident pulse (and the small counter-moving part of the initial
    # condition) has left after 325 steps; the reflection is still inside
    steps = 340
    out = {}
    for bc in ("open", "periodic"):
        run = Em2dRun(gpu, cfg, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))
        try:
            if bc == "open":
                gpu.zpic_sim_set_open_y(C.byref(run.sim))
            w0 = field_energy(gpu, run)
            run.step(steps)
            out[bc] = field_energy(gpu, run) / w0
        finally:
            run.close()
    print(f"open_y {direction:+d}: energy left {out['open']:.3e} (bound {bound:.3e}), "
          f"periodic {out['periodic']:.6f}", flush=True)
    assert out["periodic"] > 0.5
    assert out["open"] <= bound
    assert out["open"] > 0.0


def beam_cfg(nx, ny, uy, ppc=2):
    cfg = weibel(nx, ny, ppc, uz=0.0, uth=0.0)
    sp = []
    for name, mq in (("electrons", -1.0), ("positrons", 1.0)):
        for tag, u in (("up", [0.05, uy, 0.1]), ("down", [-0.02, -0.7 * uy, 0.0])):
            sp.append({"name": f"{name}_{tag}", "m_q": mq, "ppc": [ppc, ppc], "ufl": u, "uth": [0.0] * 3,
                       "n_sort": 8})
    cfg["species"] = sp
    return cfg


def ballistic(p, u, dt, dx, dy, nx, ny, steps):
    """Cells and positions after `steps` field-free pushes, with the push's
    float32 operations (em2d/particles.c:1194-1207), and which particles
    are still inside [0, ny)."""
    f = np.float32
    ux, uy, uz = (f(v) for v in u)
    rg = f(1.0) / np.sqrt(((f(1.0) + ux * ux) + uy * uy) + uz * uz)
    mx = (f(f(dt) / dx) * rg) * ux
    my = (f(f(dt) / dy) * rg) * uy
    x, y = p["x"].astype(np.float32), p["y"].astype(np.float32)
    ix, iy = p["ix"].astype(np.int64), p["iy"].astype(np.int64)
    alive = np.ones(len(p), bool)
    for _ in range(steps):
        x1, y1 = x + mx, y + my
        di = (x1 >= 1).astype(np.int64) - (x1 < 0).astype(np.int64)
        dj = (y1 >= 1).astype(np.int64) - (y1 < 0).astype(np.int64)
        x, y = x1 - di.astype(np.float32), y1 - dj.astype(np.float32)
        ix, iy = (ix + di) % nx, iy + dj
        alive &= (iy >= 0) & (iy < ny)
    return alive, ix, iy, x, y


def test_particles_leave_exactly_where_they_cross(gpu):
    nx, ny = 24, 32
    cfg = beam_cfg(nx, ny, 0.4)
    # cell sizes as spec_new forms them: box / nx in float32
    dt = cfg["dt"]
    dx = np.float32(np.float32(cfg["box"][0]) / np.float32(nx))
    dy = np.float32(np.float32(cfg["box"][1]) / np.float32(ny))
    run = Em2dRun(gpu, cfg, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))
    try:
        gpu.zpic_sim_set_open_y(C.byref(run.sim))
        ns = len(cfg["species"])
        init = []
        for k in range(ns):
            gpu.spec_track_ids(C.byref(run.species[k]))
            init.append(A.particles_view(run.species[k]).copy())
        done = 0
        for steps in (1, 4, 15, 40, 80):
            run.step(steps - done)
            done = steps
            st = run.state()
            for f in ("E", "B", "J"):
                assert not np.any(st[f]), f"{f} should cancel exactly"
            removed = 0
            for k in range(ns):
                alive, ix, iy, x, y = ballistic(init[k], cfg["species"][k]["ufl"], dt, dx, dy, nx, ny, steps)
                ids = gpu_ids(gpu, run, k)
                got = st[f"part{k}"][np.argsort(ids)]
                assert np.array_equal(np.sort(ids), np.flatnonzero(alive)), (steps, k)
                assert np.array_equal(got["ix"], ix[alive]) and np.array_equal(got["iy"], iy[alive])
                assert np.array_equal(got["x"], x[alive]) and np.array_equal(got["y"], y[alive])
                removed += int((~alive).sum())
            print(f"open_y beam: step {steps}: {removed} particles removed", flush=True)
        assert removed > 0
    finally:
        run.close()


def test_continuity_holds_away_from_the_open_edges(gpu):
    nx, ny = 48, 40
    cfg = weibel(nx, ny, 3, uz=0.3, uth=0.3, n_sort=6, m_q2=-1.0)
    dt, dx, dy = cfg["dt"], cfg["box"][0] / nx, cfg["box"][1] / ny
    run = Em2dRun(gpu, cfg, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))

    def rho():
        return sum(run.charge(k).astype(np.float64) for k in range(2))[:ny + 1, :nx]

    try:
        gpu.zpic_sim_set_open_y(C.byref(run.sim))
        n0 = sum(int(run.species[k].np) for k in range(2))
        run.step(2)
        for _ in range(20):
            r0 = rho()
            run.step(1)
            r1 = rho()
            J = run.state()["J"].astype(np.float64)
            Jx = J[1:2 + ny, 1:1 + nx, 0]      # rows 0 .. ny
            Jy = J[0:2 + ny, 1:1 + nx, 1]      # rows -1 .. ny
            res = (r1 - r0) / dt + (Jx - np.roll(Jx, 1, axis=1)) / dx + (Jy[1:] - Jy[:-1]) / dy
            scale = np.abs((r1 - r0) / dt).max()
            # node rows 0 and ny lose the charge of particles removed there
            assert np.abs(res[1:ny]).max() <= 1e-5 * scale
        gpu.sim_sync_host(C.byref(run.sim))
        n1 = sum(int(run.species[k].np) for k in range(2))
        print(f"open_y plasma: {n0} -> {n1} particles", flush=True)
        assert n1 < n0
    finally:
        run.close()
