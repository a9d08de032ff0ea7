"""Parity of the GPU library with the oracle and the reference's golden vectors.

Particles are matched one to one through tracked ids (the GPU keeps them in
tile order, the reference in cell order).  Bars:
  * first push from a common state: every particle bit-identical;
  * cell indices (ix, iy) identical for the first 10 steps;
  * E, B and J within 1e-5 relative L2 after 100 steps (the current is
    summed in a different order on the GPU, so fields differ by FP32
    rounding that grows with the physics);
  * vacuum fields (no particles): bit-identical;
  * discrete continuity and Gauss's law at FP32 rounding.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from conftest import field_err, golden_cases, gpu_ids, load_golden, oracle_ids, rel_l2
from em2d_driver import Em2dRun, lwfa, weibel
from test_oracle import continuity_residual, filter_model, gauss_defect, yee_mode

pytestmark = pytest.mark.gpu


def gpu_run(gpu, cfg, ids=False):
    run = Em2dRun(gpu, cfg, sync=lambda r: gpu.sim_sync_host(C.byref(r.sim)))
    if ids:
        for k in range(len(cfg["species"])):
            gpu.spec_track_ids(C.byref(run.species[k]))
    return run


def oracle_run(oracle, cfg, ids=False):
    run = Em2dRun(oracle, cfg)
    if ids:
        for k in range(len(cfg["species"])):
            oracle.orc_track_ids(C.byref(run.species[k]))
    return run


def matched(gpu, oracle, g, o, k):
    """GPU and oracle particles of species k, both ordered by id."""
    pg, po = g.state()[f"part{k}"], o.state()[f"part{k}"]
    ig, io = gpu_ids(gpu, g, k), oracle_ids(oracle, o, k)
    pg, ig = pg[np.argsort(ig)], np.sort(ig)
    po, io = po[np.argsort(io)], np.sort(io)
    assert np.array_equal(ig, io), "different particle sets"
    return pg, po


def lwfa_with_ions(**kw):
    """LWFA deck plus an ion species of the same density and particles per
    cell: the charges have the same magnitude, so sim_iter pushes both
    species in one launch, with the window shift, the injected columns and
    the unsorted tail of each."""
    cfg = lwfa(**kw)
    ions = dict(cfg["species"][0], name="ions", m_q=1836.0)
    cfg["species"] = [cfg["species"][0], ions]
    return cfg


CASES = {
    "weibel": weibel(32, 32, 2),
    "weibel_2e_4ppc": weibel(48, 32, 4, n_sort=20, m_q2=-1.0),
    "smoothed": weibel(64, 48, 2, n_sort=8, smooth={"xtype": "compensated", "ytype": "binomial",
                                                    "xlevel": 2, "ylevel": 1}),
    "lwfa_moving_window": lwfa(200, 32, start=6.0),
    "lwfa_electrons_ions": lwfa_with_ions(nx=200, ny=32, start=6.0),
    "hot_fast": weibel(32, 32, 3, uz=0.9, uth=0.4, n_sort=5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_first_push_is_bit_identical(gpu, oracle, name):
    cfg = CASES[name]
    g, o = gpu_run(gpu, cfg, ids=True), oracle_run(oracle, cfg, ids=True)
    try:
        g.step(1)
        o.step(1)
        for k in range(len(cfg["species"])):
            pg, po = matched(gpu, oracle, g, o, k)
            assert pg.tobytes() == po.tobytes()
        sg, so = g.state(), o.state()
        for f in ("J", "E", "B"):
            assert field_err(sg, so, f) <= 1e-5, f
        assert np.allclose(sg["kin_energy"], so["kin_energy"], rtol=1e-9)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_cell_indices_exact_for_10_steps(gpu, oracle, name):
    cfg = CASES[name]
    g, o = gpu_run(gpu, cfg, ids=True), oracle_run(oracle, cfg, ids=True)
    try:
        for step in range(1, 11):
            g.step(1)
            o.step(1)
            for k in range(len(cfg["species"])):
                pg, po = matched(gpu, oracle, g, o, k)
                assert np.array_equal(pg["ix"], po["ix"]) and np.array_equal(pg["iy"], po["iy"]), step
        sg, so = g.state(), o.state()
        for f in ("J", "E", "B"):
            assert field_err(sg, so, f) <= 1e-5, f
        assert sg["emf_n_move"] == so["emf_n_move"]
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("cfg", [weibel(64, 64, 2, m_q2=-1.0), weibel(32, 32, 2)], ids=["2e_64", "deck_32"])
def test_fields_within_1e5_after_100_steps(gpu, oracle, cfg):
    g, o = gpu_run(gpu, cfg), oracle_run(oracle, cfg)
    try:
        g.step(100)
        o.step(100)
        sg, so = g.state(), o.state()
        for f in ("E", "B", "J"):
            assert rel_l2(sg[f], so[f]) <= 1e-5, f
        for k in range(2):
            assert len(sg[f"part{k}"]) == len(so[f"part{k}"])
        assert np.allclose(sg["emf_energy"], so["emf_energy"], rtol=1e-4)
        assert np.allclose(sg["kin_energy"], so["kin_energy"], rtol=1e-5)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("waves", [0, 2, 8])
def test_float_atomic_deposit_variants_match_oracle(gpu, oracle, waves):
    """zpic_set_fixed_deposit(0): the deposit sums the tile windows with
    float LDS atomics instead of 64-bit fixed point; each variant of that
    kernel (zpic_set_push_waves) keeps the 100-step parity bars."""
    cfg = weibel(64, 64, 2, m_q2=-1.0)
    gpu.zpic_set_fixed_deposit(0)
    gpu.zpic_set_push_waves(waves)
    try:
        g, o = gpu_run(gpu, cfg), oracle_run(oracle, cfg)
        try:
            g.step(100)
            o.step(100)
            sg, so = g.state(), o.state()
            for f in ("E", "B", "J"):
                assert rel_l2(sg[f], so[f]) <= 1e-5, f
            assert np.allclose(sg["kin_energy"], so["kin_energy"], rtol=1e-5)
        finally:
            g.close()
            o.close()
    finally:
        gpu.zpic_set_fixed_deposit(1)
        gpu.zpic_set_push_waves(0)


def test_fused_species_push_matches_per_species_launches(gpu):
    """zpic_set_species_fusion: sim_iter pushes the two species of a Weibel
    run in one tile launch.  From the same fields the first push gives the
    particles of one launch per species bit for bit; the fused launch rounds
    the summed current window once instead of once per species, so after
    50 steps fields and energies agree within FP32 rounding."""
    cfg = weibel(64, 64, 2, m_q2=-1.0)
    fused, plain = gpu_run(gpu, cfg, ids=True), gpu_run(gpu, cfg, ids=True)

    def step(n):
        for _ in range(n):
            gpu.zpic_set_species_fusion(1)
            fused.step(1)
            gpu.zpic_set_species_fusion(0)
            plain.step(1)
        gpu.zpic_set_species_fusion(1)

    try:
        step(1)
        for k in range(2):
            pf, pp = fused.state()[f"part{k}"], plain.state()[f"part{k}"]
            idf, idp = gpu_ids(gpu, fused, k), gpu_ids(gpu, plain, k)
            assert np.array_equal(np.sort(idf), np.sort(idp))
            pf, pp = pf[np.argsort(idf)], pp[np.argsort(idp)]
            for f in pf.dtype.names:
                assert np.array_equal(pf[f], pp[f]), (k, f)
        sf, sp = fused.state(), plain.state()
        assert field_err(sf, sp, "J") <= 1e-6
        step(49)
        sf, sp = fused.state(), plain.state()
        for f in ("E", "B", "J"):
            assert rel_l2(sf[f], sp[f]) <= 1e-5, f
        assert np.allclose(sf["kin_energy"], sp["kin_energy"], rtol=1e-5)
    finally:
        gpu.zpic_set_species_fusion(1)
        fused.close()
        plain.close()


@pytest.mark.parametrize("open_y", [0, 1], ids=["periodic_y", "open_y"])
def test_window_shift_in_the_field_update_matches_the_separate_shift(gpu, open_y):
    """zpic_set_window_shift_fusion: on a step that ends in a window move
    the Yee kernel writes the shifted grid (em2d/emf.c:646-676) instead of
    a copy of the whole grid after it; 40 steps of an LWFA deck (36 moves)
    give the same fields, current and particles bit for bit."""
    cfg = lwfa(200, 32)
    runs = [gpu_run(gpu, cfg, ids=True), gpu_run(gpu, cfg, ids=True)]
    try:
        if open_y:
            for r in runs:
                gpu.zpic_sim_set_open_y(C.byref(r.sim))
        for _ in range(40):
            for on, r in zip((1, 0), runs):
                gpu.zpic_set_window_shift_fusion(on)
                r.step(1)
        gpu.zpic_set_window_shift_fusion(1)
        sa, sb = runs[0].state(), runs[1].state()
        assert sa["emf_n_move"] == sb["emf_n_move"] >= 30
        for f in ("E", "B", "J"):
            assert sa[f].tobytes() == sb[f].tobytes(), f
        for k in range(len(cfg["species"])):
            ia, ib = gpu_ids(gpu, runs[0], k), gpu_ids(gpu, runs[1], k)
            pa, pb = sa[f"part{k}"][np.argsort(ia)], sb[f"part{k}"][np.argsort(ib)]
            assert np.array_equal(np.sort(ia), np.sort(ib))
            assert pa.tobytes() == pb.tobytes(), k
    finally:
        gpu.zpic_set_window_shift_fusion(1)
        for r in runs:
            r.close()


@pytest.mark.parametrize("case", golden_cases())
def test_against_reference_golden_vectors(gpu, case):
    """Directly against the reference's stored states: the particle multiset
    after the first push is bit-identical; afterwards cell occupancy per
    cell is identical up to step 10 and fields agree to 1e-5."""
    cfg, steps, z = load_golden(case)
    g = gpu_run(gpu, cfg)
    try:
        for s in steps:
            if s == 0:
                continue
            g.step(s - g.n)
            st = g.state()
            ref = {f: z[f"{s}/{f}"] for f in ("E", "B", "J")}
            for f in ("E", "B", "J"):
                err = rel_l2(st[f], ref[f]) if s >= 100 else field_err(st, ref, f)
                assert err <= 1e-5, (s, f)
            for k in range(len(cfg["species"])):
                exp = z[f"{s}/part{k}"]
                assert st[f"n_move{k}"] == int(z[f"{s}/n_move{k}"])
                if not exp.size:
                    continue
                got = st[f"part{k}"]
                assert len(got) == len(exp)
                if s == 1:
                    key = lambda p: np.sort(p.view(np.dtype((np.void, 28))))
                    assert np.array_equal(key(got), key(exp)), "first push differs from the reference"
                if s <= 10:
                    nx = cfg["nx"][0]
                    assert np.array_equal(np.bincount(got["ix"] + nx * got["iy"], minlength=nx * cfg["nx"][1]),
                                          np.bincount(exp["ix"] + nx * exp["iy"], minlength=nx * cfg["nx"][1]))
    finally:
        g.close()


@pytest.mark.parametrize("comp", ["Ez", "Bz"])
def test_vacuum_fields_bit_identical(gpu, oracle, comp):
    """Without particles the field solver is deterministic: the fused GPU Yee
    kernel reproduces the oracle bit for bit, and follows the discrete
    dispersion relation."""
    nx, ny, dx, dy, dt = 64, 40, 0.1, 0.15, 0.07
    cfg, w = yee_mode(nx, ny, dx, dy, dt, 3, 2, comp)
    g, o = gpu_run(gpu, cfg), oracle_run(oracle, cfg)
    grid = "E" if comp == "Ez" else "B"
    try:
        f0 = o.state()[grid][1:1 + ny, 1:1 + nx, 2].astype(np.float64)
        g.step(60)
        o.step(60)
        sg, so = g.state(), o.state()
        assert sg["E"].tobytes() == so["E"].tobytes()
        assert sg["B"].tobytes() == so["B"].tobytes()
        f = sg[grid][1:1 + ny, 1:1 + nx, 2].astype(np.float64)
        assert np.abs(f - math.cos(w * 60 * dt) * f0).max() < 2e-5
    finally:
        g.close()
        o.close()


def _rho(run, cfg):
    nx, ny = cfg["nx"]
    return sum(run.charge(k).astype(np.float64) for k in range(len(cfg["species"])))[:ny, :nx]


@pytest.mark.parametrize("cfg", [weibel(128, 96, 3, uth=0.3, m_q2=-1.0), weibel(64, 64, 2, uz=0.9, uth=0.5)],
                         ids=["2e_128x96", "hot_64"])
def test_gpu_deposit_conserves_charge(gpu, cfg):
    """The tiled LDS deposit plus window reduction keeps the discrete
    continuity equation at FP32 rounding, and Gauss's law defect constant."""
    g = gpu_run(gpu, cfg)
    nx, ny = cfg["nx"]
    dx, dy, dt = cfg["box"][0] / nx, cfg["box"][1] / ny, cfg["dt"]
    try:
        g.step(3)
        gd0 = gauss_defect(g.state()["E"], _rho(g, cfg), dx, dy)
        for _ in range(25):  # spans sorts and window overflow of fast particles
            r0 = _rho(g, cfg)
            g.step(1)
            r1 = _rho(g, cfg)
            res, scale = continuity_residual(r0, r1, g.state()["J"], dt, dx, dy)
            assert res <= 1e-5 * scale
        gd1 = gauss_defect(g.state()["E"], _rho(g, cfg), dx, dy)
        # scale: one species' charge density (a neutral plasma has rho ~ 0)
        scale = np.abs(g.charge(0)).max()
        assert np.abs(gd1 - gd0).max() <= 1e-5 * scale
    finally:
        g.close()


def test_gpu_smoothing_matches_filter_model(gpu):
    base = weibel(64, 48, 2)
    sm = weibel(64, 48, 2, smooth={"xtype": "compensated", "ytype": "binomial", "xlevel": 2, "ylevel": 1})
    a, b = gpu_run(gpu, base), gpu_run(gpu, sm)
    try:
        a.step(1)
        b.step(1)
        Jr = a.state()["J"].astype(np.float64)
        Js = b.state()["J"].astype(np.float64)
        n = 2
        t = -2.0 + (4.0 + 2.0 * n) / n
        xp = [(0.25, 0.5)] * n + [(-1.0 / t, ((4.0 + 2.0 * n) / n) / t)]
        exp = filter_model(Jr, xp, [(0.25, 0.5)] * n)
        assert np.abs(Js - exp)[1:49, 1:65].max() <= 2e-6 * np.abs(Jr).max()
        # the deposit is order-free, so both runs fold the same J; the passes
        # are then the FP32 expression of the reference bit for bit, with the
        # compensator's weights as the host computes them (float b, total)
        f32 = np.float32
        wb = f32((4.0 + 2.0 * n) / n)
        tot = f32(2) * f32(-1) + wb
        xp32 = [(0.25, 0.5)] * n + [(f32(-1) / tot, wb / tot)]
        exp32 = filter_model(a.state()["J"], xp32, [(0.25, 0.5)] * n, dtype=f32)
        assert Js.astype(f32)[1:49, 1:65].tobytes() == exp32[1:49, 1:65].tobytes()
    finally:
        a.close()
        b.close()


def test_gpu_one_x_one_y_pass_matches_filter_model(gpu):
    """One binomial pass along x and one along y (the 2048^2 benchmark's
    filter) run as one launch (smooth_xy_kernel): the filtered J is the FP32
    model of the reference's two passes bit for bit, with the y pass reading
    the x pass's values on interior rows and the unfiltered guard rows."""
    base = weibel(64, 48, 2)
    sm = weibel(64, 48, 2, smooth={"xtype": "binomial", "ytype": "binomial", "xlevel": 1, "ylevel": 1})
    a, b = gpu_run(gpu, base), gpu_run(gpu, sm)
    try:
        a.step(1)
        b.step(1)
        f32 = np.float32
        exp32 = filter_model(a.state()["J"], [(0.25, 0.5)], [(0.25, 0.5)], dtype=f32)
        Js = b.state()["J"].astype(f32)
        assert Js[1:49, 1:65].tobytes() == exp32[1:49, 1:65].tobytes()
    finally:
        a.close()
        b.close()


def test_external_and_custom_fields(gpu, oracle):
    cfg = weibel(32, 32, 2)
    cfg["ext_fld"] = {"E_0": [0.01, 0.0, 0.0],
                      "B_custom": lambda i, dx, j, dy: (0.0, 0.0, 0.3 + 0.1 * math.sin(i * dx + j * dy))}
    cfg["init_fld"] = {"E_custom": lambda i, dx, j, dy: (0.0, 0.02 * math.cos(j * dy), 0.0)}
    g, o = gpu_run(gpu, cfg, ids=True), oracle_run(oracle, cfg, ids=True)
    try:
        g.step(1)
        o.step(1)
        for k in range(2):
            pg, po = matched(gpu, oracle, g, o, k)
            assert pg.tobytes() == po.tobytes()
        g.step(9)
        o.step(9)
        sg, so = g.state(), o.state()
        for f in ("E", "B", "J"):
            assert field_err(sg, so, f) <= 1e-5
        assert np.all(np.isfinite(A_grid(g)))
    finally:
        g.close()
        o.close()


def A_grid(run):
    from zpic_amd import abi as A

    return A.grid_view(run.sim.emf.ext_fld.B_part_buf, tuple(run.cfg["nx"]))


def test_charge_and_energies_match_oracle(gpu, oracle):
    cfg = weibel(48, 40, 3, n_sort=7)
    g, o = gpu_run(gpu, cfg), oracle_run(oracle, cfg)
    try:
        g.step(12)
        o.step(12)
        for k in range(2):
            assert rel_l2(g.charge(k), o.charge(k)) <= 1e-5
        sg, so = g.state(), o.state()
        assert np.allclose(sg["emf_energy"], so["emf_energy"], rtol=1e-4)
        assert np.allclose(sg["kin_energy"], so["kin_energy"], rtol=1e-6)
    finally:
        g.close()
        o.close()


def test_push_counter_and_timers(gpu):
    before = gpu.spec_npush()
    cfg = weibel(32, 32, 2)
    g = gpu_run(gpu, cfg)
    try:
        g.step(7)
        gpu.zpic_synchronize()
        assert gpu.spec_npush() - before == 7 * 2 * 32 * 32 * 4
        assert gpu.spec_time() > 0 and gpu.emf_time() > 0
    finally:
        g.close()


def test_host_edits_are_uploaded(gpu, oracle):
    """Editing the host mirror (what the Python `particles` view allows)
    changes the next step exactly as it changes the reference."""
    cfg = weibel(16, 16, 2)
    g, o = gpu_run(gpu, cfg), oracle_run(oracle, cfg)
    try:
        g.step(2)
        o.step(2)
        st = g.state()  # synced mirror, in GPU order
        for run, lib in ((g, gpu), (o, oracle)):
            sp = run.species[0]
            for i in range(sp.np):
                sp.part[i].ux = 0.3
        gpu.spec_mark_host_dirty(C.byref(g.species[0]))
        g.step(1)
        o.step(1)
        sg, so = g.state(), o.state()
        assert len(st["part0"]) == len(sg["part0"])
        assert rel_l2(sg["J"], so["J"]) <= 1e-5
        assert np.allclose(np.sort(sg["part0"]["ux"]), np.sort(so["part0"]["ux"]), rtol=1e-5, atol=1e-7)
    finally:
        g.close()
        o.close()
