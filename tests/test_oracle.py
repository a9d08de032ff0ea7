"""The oracle, pinned: against the reference's own golden vectors, against the
compiled reference on fresh inputs, and against what the mathematics fixes
(discrete continuity, Gauss's law, the Yee dispersion relation, Boris
gyration and E x B drift, filter weights, stable sorting, energy sums).

These tests run on the CPU.  Each check names the oracle function (and the
reference lines it follows) that it pins.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import ref_worker
from conftest import HERE, golden_cases, load_golden, oracle_ids
from em2d_driver import Em2dRun, weibel


# --------------------------------------------------------------------------
# golden vectors written by the reference (tests/golden/make_golden.py)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("case", golden_cases())
def test_oracle_reproduces_reference_golden_vectors(oracle, case):
    """Whole-step parity, bit for bit: injection (particles.c:85-492), push
    and deposit (:773-1259), sort (:942-1007), current fold and smoothing
    (current.c:118-459), Yee solver and guards (emf.c:500-716), window
    (particles.c:619-640, emf.c:648-675), energies (emf.c:729-750)."""
    cfg, steps, z = load_golden(case)
    run = Em2dRun(oracle, cfg)
    try:
        for s in steps:
            run.step(s - run.n)
            st = run.state()
            for key in ("E", "B", "J", "emf_energy", "kin_energy"):
                exp = z[f"{s}/{key}"]
                assert np.asarray(st[key]).tobytes() == exp.tobytes(), f"{case} step {s}: {key} differs"
            for k in range(len(cfg["species"])):
                exp = z[f"{s}/part{k}"]
                if exp.size:
                    assert st[f"part{k}"].tobytes() == exp.tobytes(), f"{case} step {s}: particles of {k}"
                assert st[f"n_move{k}"] == int(z[f"{s}/n_move{k}"])
    finally:
        run.close()


# --------------------------------------------------------------------------
# the compiled reference on inputs the golden files do not cover
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", sorted(ref_worker.cases()))
def test_oracle_matches_compiled_reference(reference, name):
    """Slab and custom densities, custom initial/external fields, a plane-wave
    laser with moving window, an odd-sized grid with mixed filters: bit-exact
    against the reference build, each in a fresh process."""
    out = subprocess.run([sys.executable, os.path.join(HERE, "ref_worker.py"), name], check=True,
                         capture_output=True, text=True, timeout=300)
    result = json.loads(out.stdout.strip().splitlines()[-1])
    assert result["mismatch"] == [], result


# --------------------------------------------------------------------------
# invariants and closed forms
# --------------------------------------------------------------------------
def _rho(run, nx, ny):
    return sum(run.charge(k).astype(np.float64) for k in range(len(run.cfg["species"])))[:ny, :nx]


def continuity_residual(rho0, rho1, J, dt, dx, dy):
    """max |(rho1-rho0)/dt + dJx/dx + dJy/dy| over the periodic grid, and the
    scale of the terms.  Jx lives at (i+1/2, j), Jy at (i, j+1/2), rho at
    the nodes (particles.c:912-921, 1304-1307)."""
    ny, nx = rho0.shape
    Jx = J[1:1 + ny, 1:1 + nx, 0].astype(np.float64)
    Jy = J[1:1 + ny, 1:1 + nx, 1].astype(np.float64)
    res = (rho1 - rho0) / dt + (Jx - np.roll(Jx, 1, axis=1)) / dx + (Jy - np.roll(Jy, 1, axis=0)) / dy
    return float(np.abs(res).max()), float(np.abs((rho1 - rho0) / dt).max())


def test_deposit_satisfies_discrete_continuity(oracle):
    """dep_current_zamb (particles.c:773-924) with the periodic fold
    (current.c:118-159) conserves charge exactly: a wrong weight, sign or
    segment cut leaves an O(1) residual."""
    cfg = weibel(32, 24, 2, uth=0.3)
    run = Em2dRun(oracle, cfg)
    dx, dy, dt = cfg["box"][0] / 32, cfg["box"][1] / 24, cfg["dt"]
    try:
        run.step(3)
        for _ in range(3):
            r0 = _rho(run, 32, 24)
            run.step()
            r1 = _rho(run, 32, 24)
            res, scale = continuity_residual(r0, r1, run.state()["J"], dt, dx, dy)
            assert res <= 1e-5 * scale
    finally:
        run.close()


def gauss_defect(E, rho, dx, dy):
    ny, nx = rho.shape
    Ex = E[1:1 + ny, 1:1 + nx, 0].astype(np.float64)
    Ey = E[1:1 + ny, 1:1 + nx, 1].astype(np.float64)
    return (Ex - np.roll(Ex, 1, axis=1)) / dx + (Ey - np.roll(Ey, 1, axis=0)) / dy - rho


def test_gauss_law_defect_is_constant(oracle):
    """Yee update (emf.c:531-562) + charge-conserving deposit keep
    div E - rho constant in time."""
    cfg = weibel(32, 32, 2, uth=0.2, m_q2=-1.0)
    run = Em2dRun(oracle, cfg)
    dx, dy = cfg["box"][0] / 32, cfg["box"][1] / 32
    try:
        run.step(1)
        g1 = gauss_defect(run.state()["E"], _rho(run, 32, 32), dx, dy)
        run.step(30)
        g2 = gauss_defect(run.state()["E"], _rho(run, 32, 32), dx, dy)
        assert np.abs(g2 - g1).max() <= 1e-5 * np.abs(run.charge(0)).max()
    finally:
        run.close()


def yee_mode(nx, ny, dx, dy, dt, m, n, comp):
    """Vacuum standing mode and its discrete frequency:
    sin^2(w dt/2)/dt^2 = sin^2(kx dx/2)/dx^2 + sin^2(ky dy/2)/dy^2."""
    kx, ky = 2 * math.pi * m / (nx * dx), 2 * math.pi * n / (ny * dy)
    s = math.sqrt((math.sin(kx * dx / 2) / dx) ** 2 + (math.sin(ky * dy / 2) / dy) ** 2)
    w = 2 * math.asin(dt * s) / dt
    if comp == "Ez":  # E_z at the nodes
        fn = lambda i, hx, j, hy: (0.0, 0.0, math.cos(kx * i * dx + ky * j * dy))
        key = "E_custom"
    else:  # B_z at the cell centres
        fn = lambda i, hx, j, hy: (0.0, 0.0, math.cos(kx * (i + 0.5) * dx + ky * (j + 0.5) * dy))
        key = "B_custom"
    cfg = {"nx": [nx, ny], "box": [nx * dx, ny * dy], "dt": dt, "species": [], "init_fld": {key: fn}}
    return cfg, w


@pytest.mark.parametrize("comp", ["Ez", "Bz"])
def test_yee_vacuum_mode_follows_discrete_dispersion(oracle, comp):
    """yee_b / yee_e (emf.c:500-562) and emf_update_gc (:573-637): a vacuum
    standing wave keeps its shape and oscillates as cos(w n dt) with the
    discrete Yee frequency."""
    nx, ny, dx, dy, dt = 32, 16, 0.1, 0.15, 0.07
    cfg, w = yee_mode(nx, ny, dx, dy, dt, 3, 2, comp)
    run = Em2dRun(oracle, cfg)
    grid = "E" if comp == "Ez" else "B"
    try:
        f0 = run.state()[grid][1:1 + ny, 1:1 + nx, 2].astype(np.float64)
        for n in (1, 17, 60):
            run.step(n - run.n)
            f = run.state()[grid][1:1 + ny, 1:1 + nx, 2].astype(np.float64)
            assert np.abs(f - math.cos(w * n * dt) * f0).max() < 2e-5
    finally:
        run.close()


def _sparse(nx, ppc, ufl, init_fld, dt=0.07, dx=0.1):
    """One nearly chargeless cold species, so self-fields are negligible; no
    sorting, so buffer positions keep their particles."""
    return {"nx": [nx, nx], "box": [nx * dx, nx * dx], "dt": dt,
            "species": [{"name": "probe", "m_q": -1.0, "ppc": [ppc, ppc], "ufl": list(ufl),
                         "uth": [0, 0, 0], "n_sort": 0, "density": {"type": "uniform", "n": 1e-9}}],
            "init_fld": init_fld}


def test_boris_rotation_angle_and_speed(oracle):
    """Boris push (particles.c:1137-1192) in a uniform B_z: |u| is kept and
    u turns by 2*atan(B dt / (2 gamma)) per step, counter-clockwise for
    q/m = -1."""
    B, u0, steps = 1.0, 0.1, 50
    cfg = _sparse(8, 1, [u0, 0, 0], {"B_0": [0.0, 0.0, B]})
    run = Em2dRun(oracle, cfg)
    try:
        run.step(steps)
        p = run.state()["part0"]
        g = math.sqrt(1 + u0 * u0)
        theta = 2 * math.atan(B * cfg["dt"] / (2 * g)) * steps
        u = p["ux"].astype(np.float64) + 1j * p["uy"].astype(np.float64)
        assert np.abs(np.abs(u) - u0).max() < 1e-6
        assert np.abs(u - u0 * np.exp(1j * theta)).max() < 1e-5
    finally:
        run.close()


def test_exb_drift(oracle):
    """Gather + Boris + move: particles at rest in E_y and B_z drift along x
    at E_y/B_z (a dropped or mis-signed term breaks the drift)."""
    E, B = 0.01, 1.0
    period = 2 * math.pi / B
    cfg = _sparse(8, 1, [0, 0, 0], {"E_0": [0.0, E, 0.0], "B_0": [0.0, 0.0, B]})
    steps = int(round(3 * period / cfg["dt"]))
    run = Em2dRun(oracle, cfg)
    try:
        p0 = run.state()["part0"]
        run.step(steps)
        p1 = run.state()["part0"]
        x0 = (p0["ix"] + p0["x"]) * 0.1
        x1 = (p1["ix"] + p1["x"]) * 0.1
        dxp = np.mod(x1 - x0 + 0.4, 0.8) - 0.4
        drift = dxp.mean() / (steps * cfg["dt"])
        assert abs(drift - E / B) < 0.03 * (E / B)
    finally:
        run.close()


def test_uniform_field_is_interpolated_exactly(oracle):
    """interpolate_fld (particles.c:1029-1071) reproduces a uniform field:
    one push of particles at rest gives u = (q/m) E dt."""
    E = 0.02
    cfg = _sparse(8, 2, [0, 0, 0], {"E_0": [E, -E, 0.5 * E]})
    run = Em2dRun(oracle, cfg)
    try:
        run.step(1)
        p = run.state()["part0"]
        dt = cfg["dt"]
        for f, e in (("ux", E), ("uy", -E), ("uz", 0.5 * E)):
            assert np.allclose(p[f], -e * dt, rtol=2e-6, atol=0)
    finally:
        run.close()


def _binomial(J, axis, passes, sa=0.25, sb=0.5):
    for _ in range(passes):
        J = sa * np.roll(J, 1, axis=axis) + sb * J + sa * np.roll(J, -1, axis=axis)
    return J


def filter_model(J, xpasses, ypasses, dtype=np.float64):
    """Filter passes on a guard-inclusive periodic grid J[rows, cols, 3]
    (one lower, two upper guards) with the guard rules of kernel_x/kernel_y
    (current.c:316-413): an x pass refreshes the x guards of interior rows
    only, so the guard rows a following y pass reads are not x-filtered.
    With dtype float32 every product and sum is rounded to FP32 in the
    order sa*l + sb*m + sa*u, the reference's expression without
    contraction."""
    J = J.astype(dtype).copy()
    xpasses = [(dtype(sa), dtype(sb)) for sa, sb in xpasses]
    ypasses = [(dtype(sa), dtype(sb)) for sa, sb in ypasses]
    ny, nx = J.shape[0] - 3, J.shape[1] - 3
    for sa, sb in xpasses:
        I = J[1:1 + ny]
        new = sa * I[:, 0:nx] + sb * I[:, 1:1 + nx] + sa * I[:, 2:2 + nx]
        I[:, 1:1 + nx] = new
        I[:, 0] = new[:, nx - 1]
        I[:, 1 + nx] = new[:, 0]
        I[:, 2 + nx] = new[:, 1]
    for sa, sb in ypasses:
        new = sa * J[0:ny, 1:1 + nx] + sb * J[1:1 + ny, 1:1 + nx] + sa * J[2:2 + ny, 1:1 + nx]
        J[1:1 + ny, 1:1 + nx] = new
        J[0] = J[ny]
        J[1 + ny] = J[1]
        J[2 + ny] = J[2]
    return J


def test_smoothing_matches_three_point_filter_model(oracle):
    """current_smooth (current.c:427-459): binomial [1/4, 1/2, 1/4] passes in
    x then y on the folded current, the y pass count taken from xlevel
    (current.c:449).  x-only and y-only filtering are plain periodic
    convolutions; together, the y pass reads x-unfiltered guard rows."""
    base = weibel(32, 24, 2)
    for smooth, xp, yp in (({"xtype": "binomial", "xlevel": 2}, 2, 0),
                           ({"ytype": "binomial", "xlevel": 1, "ylevel": 1}, 0, 1),
                           ({"xtype": "binomial", "ytype": "binomial", "xlevel": 2, "ylevel": 1}, 2, 2)):
        a, b = Em2dRun(oracle, base), Em2dRun(oracle, weibel(32, 24, 2, smooth=smooth))
        try:
            a.step(1)
            b.step(1)
            Jr = a.state()["J"].astype(np.float64)
            Js = b.state()["J"].astype(np.float64)
            exp = filter_model(Jr, [(0.25, 0.5)] * xp, [(0.25, 0.5)] * yp)
            assert np.abs(Js - exp)[1:25, 1:33].max() <= 1e-6 * np.abs(Jr).max()
            # in FP32, operation by operation, the model is the oracle bit for bit
            exp32 = filter_model(a.state()["J"], [(0.25, 0.5)] * xp, [(0.25, 0.5)] * yp, dtype=np.float32)
            assert b.state()["J"][1:25, 1:33].tobytes() == exp32[1:25, 1:33].tobytes()
            if yp == 0:
                conv = _binomial(Jr[1:25, 1:33], 1, xp)
                assert np.abs(Js[1:25, 1:33] - conv).max() <= 1e-6 * np.abs(Jr).max()
            if xp == 0:
                conv = _binomial(Jr[1:25, 1:33], 0, yp)
                assert np.abs(Js[1:25, 1:33] - conv).max() <= 1e-6 * np.abs(Jr).max()
        finally:
            a.close()
            b.close()


def test_compensator_weights_flatten_transfer_function():
    """get_smooth_comp (current.c:297-304): after n binomial passes the
    [a, b, a] compensator with b = (4 + 2n)/n, a = -1 (normalised) cancels
    the k^2 term of the transfer function."""
    for n in (1, 2, 4):
        a, b = -1.0, (4.0 + 2.0 * n) / n
        t = 2 * a + b
        sa, sb = a / t, b / t
        for k in (1e-2, 2e-2):
            T = math.cos(k / 2) ** (2 * n) * (sb + 2 * sa * math.cos(k))
            assert abs(1 - T) < 5 * k ** 4


def test_compensated_smoothing_matches_convolution(oracle):
    base = weibel(32, 32, 2)
    comp = weibel(32, 32, 2, smooth={"xtype": "compensated", "xlevel": 3})
    a, b = Em2dRun(oracle, base), Em2dRun(oracle, comp)
    try:
        a.step(1)
        b.step(1)
        Jr = a.state()["J"][1:33, 1:33].astype(np.float64)
        Js = b.state()["J"][1:33, 1:33].astype(np.float64)
        n = 3
        t = -2.0 + (4.0 + 2.0 * n) / n
        exp = _binomial(_binomial(Jr, 1, n), 1, 1, -1.0 / t, ((4.0 + 2.0 * n) / n) / t)
        assert np.abs(Js - exp).max() <= 2e-6 * np.abs(Jr).max()
    finally:
        a.close()
        b.close()


def test_sort_is_stable_counting_sort_by_cell(oracle):
    """spec_sort (particles.c:942-1007) equals numpy's stable argsort on
    ix + iy*nx applied to the unsorted buffer of the same step."""
    sorted_run = Em2dRun(oracle, weibel(16, 16, 2, n_sort=4))
    plain_run = Em2dRun(oracle, weibel(16, 16, 2, n_sort=0))
    try:
        sorted_run.step(4)
        plain_run.step(4)
        ps = sorted_run.state()["part0"]
        pp = plain_run.state()["part0"]
        key = pp["ix"] + pp["iy"] * 16
        assert ps.tobytes() == pp[np.argsort(key, kind="stable")].tobytes()
    finally:
        sorted_run.close()
        plain_run.close()


def test_field_energy_is_sum_of_squares(oracle):
    """emf_get_energy (emf.c:729-750): 0.5*dx*dy*sum(F_c^2) per component."""
    cfg = weibel(24, 16, 2)
    run = Em2dRun(oracle, cfg)
    try:
        run.step(5)
        st = run.state()
        dx, dy = cfg["box"][0] / 24, cfg["box"][1] / 16
        for g, off in (("E", 0), ("B", 3)):
            F = st[g][1:17, 1:25].astype(np.float64)
            for c in range(3):
                assert math.isclose(st["emf_energy"][off + c], 0.5 * dx * dy * np.sum(F[..., c] ** 2),
                                    rel_tol=1e-6)
    finally:
        run.close()


def test_total_charge_and_particle_count_conserved(oracle):
    """spec_deposit_charge (particles.c:1289-1324) sums to q*np on the
    periodic grid; periodic boundaries keep every particle."""
    cfg = weibel(16, 16, 2)
    run = Em2dRun(oracle, cfg)
    try:
        q = [run.species[k].q for k in range(2)]
        run.step(20)
        for k in range(2):
            assert run.species[k].np == 16 * 16 * 4
            tot = run.charge(k)[:16, :16].astype(np.float64).sum()
            assert math.isclose(tot, q[k] * run.species[k].np, rel_tol=1e-5)
    finally:
        run.close()


def test_oracle_ids_follow_particles(oracle):
    """Test extension: ids move with the particles through sorts."""
    run = Em2dRun(oracle, weibel(16, 16, 2, n_sort=2))
    try:
        oracle.orc_track_ids(C.byref(run.species[0]))
        p0 = run.state()["part0"]
        run.step(2)
        ids = oracle_ids(oracle, run, 0)
        assert sorted(ids.tolist()) == list(range(len(p0)))
        p = run.state()["part0"]
        # a particle moves at most one cell per step
        d = np.abs(p["ix"].astype(int) - p0["ix"][ids].astype(int))
        assert (np.minimum(d, 16 - d) <= 2).all()
    finally:
        run.close()


@pytest.mark.parametrize("nx,ny", [(16, 32), (32, 16), (24, 24)])
def test_initial_momenta_remove_the_mean_per_stride_key(oracle, nx, ny):
    """spec_set_u removes the mean thermal momentum of the particles that
    share the key ix + nx[1] * iy (em2d/particles.c:103-142): the plain
    per-cell mean for square grids and for nx[1] > nx[0] (where the keys are
    distinct; the reference's arrays are too small there), aliased cells
    for nx[0] > nx[1].  Checked by grouping the initial particles by key."""
    run = Em2dRun(oracle, weibel(nx, ny, 3, uz=0.6, uth=0.1))
    try:
        p = run.state()["part0"]
        key = p["ix"].astype(np.int64) + ny * p["iy"].astype(np.int64)
        for comp, fluid in (("ux", 0.0), ("uy", 0.0), ("uz", 0.6)):
            th = p[comp].astype(np.float64) - fluid
            sums = np.bincount(key, weights=th)
            cnt = np.bincount(key)
            used = cnt > 0
            assert np.abs(sums[used] / cnt[used]).max() < 2e-6, comp
        # the group sizes tell the aliasing apart from plain cells
        sizes = np.unique(np.bincount(key)[np.bincount(key) > 0])
        expect = {9} if nx <= ny else set(sizes.tolist())
        assert set(sizes.tolist()) == expect
        if nx > ny:
            assert sizes.max() > 9  # some keys gather several cells
    finally:
        run.close()
