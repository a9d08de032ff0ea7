"""Batched simulations (zpic_batch_*): every simulation of a batch evolves
as the same deck does on its own.

The batch has no counterpart in the reference; each member is pinned by
the oracle run of the same deck, with the bars of the single-simulation
parity tests: cell indices identical, particles matched through ids, E, B
and J within 1e-5 relative, energies within FP32 rounding.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from conftest import field_err, gpu_ids, oracle_ids
from em2d_driver import Em2dRun, weibel

pytestmark = pytest.mark.gpu


def decks():
    """Members of uneven size: grids, particle counts, species counts,
    filters and sort periods all differ; one member is a vacuum run."""
    three = weibel(32, 64, 2, n_sort=4, seed=(7, 11))
    three["species"].append({"name": "ions", "m_q": 50.0, "ppc": [1, 2], "ufl": [0.1, 0, 0],
                             "uth": [0.01, 0.01, 0.01], "n_sort": 4})
    one = weibel(48, 48, 3, n_sort=0, seed=(5, 9))
    one["species"] = one["species"][:1]
    mode = 2 * math.pi / 32
    vacuum = {"nx": [32, 32], "box": [3.2, 3.2], "dt": 0.07, "species": [],
              "init_fld": {"B_custom": lambda ix, dx, iy, dy: (0.0, 0.0, 1e-3 * math.sin(mode * (ix + dx)))}}
    return [
        weibel(32, 32, 2, n_sort=5),
        weibel(64, 32, 3, n_sort=7, seed=(1, 2),
               smooth={"xtype": "compensated", "ytype": "binomial", "xlevel": 2, "ylevel": 1}),
        one,
        three,
        weibel(48, 64, 4, uz=0.3, uth=0.05, n_sort=20, m_q2=-1.0, seed=(3, 4),
               smooth={"xtype": "binomial", "ytype": "binomial", "xlevel": 1, "ylevel": 1}),
        vacuum,
    ]


def make(lib, cfg, track, kind):
    run = Em2dRun(lib, cfg, sync=(lambda r: lib.sim_sync_host(C.byref(r.sim))) if kind == "gpu" else None)
    for k in range(len(cfg["species"])):
        track(C.byref(run.species[k]))
    return run


def new_batch(gpu, runs):
    arr = (C.c_void_p * len(runs))(*[C.addressof(r.sim) for r in runs])
    b = gpu.zpic_batch_new(arr, len(runs))
    assert b
    return b


def compare(gpu, oracle, g, o, cfg, ftol=1e-5):
    sg, so = g.state(), o.state()
    for k in range(len(cfg["species"])):
        ig, io = gpu_ids(gpu, g, k), oracle_ids(oracle, o, k)
        pg, po = sg[f"part{k}"][np.argsort(ig)], so[f"part{k}"][np.argsort(io)]
        assert np.array_equal(np.sort(ig), np.sort(io))
        assert np.array_equal(pg["ix"], po["ix"]) and np.array_equal(pg["iy"], po["iy"])
        for f in ("x", "y", "ux", "uy", "uz"):
            assert np.allclose(pg[f], po[f], rtol=0, atol=1e-4), f
    for f in ("E", "B", "J"):
        assert field_err(sg, so, f) <= ftol, f
    assert np.allclose(sg["kin_energy"], so["kin_energy"], rtol=1e-5, atol=0)
    assert np.allclose(sg["emf_energy"], so["emf_energy"], rtol=1e-4, atol=1e-12)


@pytest.fixture
def batch_api(gpu):
    gpu.zpic_batch_new.restype = C.c_void_p
    gpu.zpic_batch_new.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    gpu.zpic_batch_run.argtypes = [C.c_void_p, C.c_int]
    gpu.zpic_batch_delete.argtypes = [C.c_void_p]
    return gpu


def test_batch_members_match_their_own_runs(batch_api, oracle):
    gpu = batch_api
    cfgs = decks()
    gs = [make(gpu, c, gpu.spec_track_ids, "gpu") for c in cfgs]
    b = new_batch(gpu, gs)
    try:
        print("batch: 6 members, 12 steps", flush=True)
        gpu.zpic_batch_run(b, 12)
        for i, cfg in enumerate(cfgs):
            o = make(oracle, cfg, oracle.orc_track_ids, "oracle")
            try:
                o.step(12)
                compare(gpu, oracle, gs[i], o, cfg)
            finally:
                o.close()
    finally:
        gpu.zpic_batch_delete(b)
        for g in gs:
            g.close()


def test_batch_runs_interleave_with_single_steps(batch_api, oracle):
    """batch run, single steps, batch run again: the host-side state
    (iteration counts, buffer parities, sort schedule) follows the device."""
    gpu = batch_api
    cfgs = decks()[:4]
    gs = [make(gpu, c, gpu.spec_track_ids, "gpu") for c in cfgs]
    b = new_batch(gpu, gs)
    try:
        print("batch: 7 batched + 3 single + 9 batched steps", flush=True)
        gpu.zpic_batch_run(b, 7)
        for g in gs:
            g.step(3)
        gs[1].state()  # a host sync between runs
        gpu.zpic_batch_run(b, 9)
        for i, cfg in enumerate(cfgs):
            o = make(oracle, cfg, oracle.orc_track_ids, "oracle")
            try:
                o.step(19)
                compare(gpu, oracle, gs[i], o, cfg)
            finally:
                o.close()
    finally:
        gpu.zpic_batch_delete(b)
        for g in gs:
            g.close()


def test_batch_matches_single_gpu_runs_over_many_steps(batch_api):
    """Sixty steps of a batch against the same decks stepped one by one on
    the GPU: cell indices identical, fields within 1e-5."""
    gpu = batch_api
    cfgs = decks()
    gs = [make(gpu, c, gpu.spec_track_ids, "gpu") for c in cfgs]
    b = new_batch(gpu, gs)
    try:
        print("batch: 60 steps vs single runs", flush=True)
        gpu.zpic_batch_run(b, 60)
        for i, cfg in enumerate(cfgs):
            s = make(gpu, cfg, gpu.spec_track_ids, "gpu")
            try:
                s.step(60)
                sg, ss = gs[i].state(), s.state()
                for k in range(len(cfg["species"])):
                    ig, is_ = gpu_ids(gpu, gs[i], k), gpu_ids(gpu, s, k)
                    pg, ps = sg[f"part{k}"][np.argsort(ig)], ss[f"part{k}"][np.argsort(is_)]
                    assert np.array_equal(pg["ix"], ps["ix"]) and np.array_equal(pg["iy"], ps["iy"])
                for f in ("E", "B", "J"):
                    assert field_err(sg, ss, f) <= 1e-5, f
                assert gs[i].sim.emf.iter == s.sim.emf.iter == 60
                for k in range(len(cfg["species"])):
                    assert gs[i].species[k].iter == 60
            finally:
                s.close()
    finally:
        gpu.zpic_batch_delete(b)
        for g in gs:
            g.close()


def test_device_initialised_members_match_host_initialised(batch_api):
    """Scan members built on the device (scan.build_members(device=True))
    hold the particles the host initialisation makes, so their batch runs
    give the same fields and particles, bit for bit."""
    from zpic_amd import scan

    gpu = batch_api
    seeds = [3, 8, 21]
    params = [dict(scan.member_params(s), nx=64, ny=32 + 32 * (k % 2), ppc=2 + k) for k, s in enumerate(seeds)]
    out = []
    for device in (False, True):
        prev = gpu.zpic_device_init()
        gpu.zpic_set_device_init(1 if device else 0)
        try:
            ms = [scan.Member(gpu, p) for p in params]
        finally:
            gpu.zpic_set_device_init(prev)
        b = scan.Batch(ms)
        try:
            print(f"batch of 3 members, {'device' if device else 'host'} initialisation: 25 steps", flush=True)
            b.run(25)
            res = []
            for m in ms:
                gpu.sim_sync_host(C.byref(m.sim))
                st = {}
                nodes = (m.sim.emf.nrow) * (m.params["ny"] + 3)
                for f, buf in (("E", m.sim.emf.E_buf), ("B", m.sim.emf.B_buf), ("J", m.sim.current.J_buf)):
                    st[f] = np.ctypeslib.as_array(C.cast(buf, C.POINTER(C.c_float)), shape=(3 * nodes,)).copy()
                parts = []
                for k in range(2):
                    sp = m.species[k]
                    raw = np.ctypeslib.as_array(C.cast(sp.part, C.POINTER(C.c_int32)), shape=(7 * sp.np,))
                    rows = raw.reshape(sp.np, 7).copy()
                    parts.append(rows[np.lexsort(rows.T[::-1])])
                st["parts"] = parts
                res.append(st)
            out.append(res)
        finally:
            b.close()
            for m in ms:
                m.close()
    host, dev = out
    for h, d in zip(host, dev):
        for f in ("E", "B", "J"):
            assert h[f].tobytes() == d[f].tobytes(), f
        for k in range(2):
            assert h["parts"][k].tobytes() == d["parts"][k].tobytes(), k
