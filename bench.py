#!/usr/bin/env python3
"""Benchmark of the em2d time step on one MI355X.

Workload (the large Weibel configuration): 2048 x 2048 periodic grid,
dx = dy = 0.1, dt = 0.07, two counter-streaming electron species (m_q = -1,
u_z = +-0.6, u_th = 0.1), 4 x 4 particles per cell each (134,217,728
particles), tile sort every 20 steps, one binomial smoothing pass along x
and y.  The default run is the full 1000-step simulation: 10 untimed
warm-up steps, then 990 timed steps.

Prints one JSON line: particle-steps/s over the timed steps, the HBM
roofline of the dominant kernel (push + deposit, timed with HIP events on
the library's stream), and a CPU baseline: the reference em2d built with
its own flags (oracle/_ref/ref_bench), run on the host cores over a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def ensure_built() -> None:
    if not os.path.exists(os.path.join(ROOT, "zpic_amd", "libzpic_em2d.so")):
        import __graft_entry__

        __graft_entry__.build()


def host_cores() -> dict:
    """Cores of the host and of this process's share: the affinity mask,
    capped at 16, the host cores allotted to one GPU on the GPU boxes."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"host_cpu_count": os.cpu_count(), "affinity": aff, "share": min(16, aff)}


def cpu_baseline(sample_nx: int, ppc: int, steps: int, nproc: int) -> dict | None:
    """Reference em2d (its own -Ofast build) on `nproc` host cores at once,
    each running the sample; aggregate particle-steps/s."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    kind = "reference"
    if not os.path.exists(exe):
        exe = os.path.join(ROOT, "oracle", "_build", "oracle_bench")
        kind = "port"
    if not os.path.exists(exe):
        return None
    args = [exe, str(sample_nx), str(sample_nx), str(ppc), str(ppc), str(steps), "20", "1", "0.1", "0.07", "0.6",
            "0.1", "-1"]
    t0 = time.time()
    procs = [subprocess.Popen(args, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for _ in range(nproc)]
    outs = [p.communicate()[0] for p in procs]
    wall = time.time() - t0
    rates = []
    for o in outs:
        try:
            rates.append(json.loads(o.strip().splitlines()[-1])["particle_steps_per_s"])
        except Exception:
            return None
    npart = 2 * sample_nx * sample_nx * ppc * ppc
    hc = host_cores()
    return {"value": float(sum(rates)), "unit": "particle-steps/s", "cores": nproc, "kind": kind,
            "single_core_value": float(sorted(rates)[len(rates) // 2]),
            "host_cpu_count": hc["host_cpu_count"], "affinity_cores": hc["affinity"],
            "sample": (f"same physics on a {sample_nx}x{sample_nx} grid ({npart} particles), {steps} steps "
                       f"(tile sorts every 20 steps: {steps // 20} sorts), {nproc} independent processes on "
                       f"{nproc} host cores (the GPU's share of the host: {hc['affinity']} cores in the "
                       f"affinity mask, {hc['host_cpu_count']} on the host), {wall:.1f} s wall")}


def copy_probe(nbytes: int) -> int:
    """One hipMemcpy device-to-device of nbytes: the runtime's copy kernel
    reads and writes exactly nbytes, the yardstick for HBM byte counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": nbytes}))
    return 0


def member_energy(m) -> float:
    """Total field + kinetic energy of a scan member (sim_report_energy's sum)."""
    import ctypes as C

    lib = m.lib
    lib.sim_sync_host(C.byref(m.sim))
    e6 = (C.c_double * 6)()
    lib.emf_get_energy(C.byref(m.sim.emf), e6)
    return float(sum(e6[:]) + m.species[0].energy + m.species[1].energy)


def batch_workload(args) -> dict:
    """Parameter scan of `--members` two-stream / Weibel simulations of
    uneven size (zpic_amd.scan), `--batch-steps` steps each: one batched
    run on the GPU against the same members run one after another with
    sim_iter (each built, run and freed in turn)."""
    import ctypes as C

    from zpic_amd import _lib
    from zpic_amd import scan

    lib = _lib.load()
    lib.zpic_set_push_waves(args.waves)
    seeds = list(range(args.members))
    steps = args.batch_steps
    threads = args.threads or min(16, os.cpu_count() or 1)
    tiles = None
    t_tune = 0.0
    if args.tune:
        t0 = time.time()
        classes = {scan.size_class(scan.member_params(s)) for s in seeds}
        log(f"batch: tuning tiles for {len(classes)} size classes")
        tiles = scan.tune_tiles(classes)
        t_tune = time.time() - t0
        log(f"batch: tiles {tiles} ({t_tune:.1f} s)")
    device = not args.host_init
    log(f"batch: building {len(seeds)} members " + ("on the device" if device else f"on {threads} host threads"))
    t0 = time.time()
    members = scan.build_members(seeds, threads=threads, tiles=tiles, device=device)
    lib.zpic_synchronize()
    t_init = time.time() - t0
    npart = sum(m.particles for m in members)
    log(f"batch: init {t_init:.1f} s, {npart} particles; setting up the batch")
    t0 = time.time()
    b = scan.Batch(members)
    lib.zpic_synchronize()
    t_setup = time.time() - t0
    log(f"batch: batch set up in {t_setup:.1f} s; running {steps} steps")
    t0 = time.perf_counter()
    for done in range(0, steps, 50):
        b.run(min(50, steps - done))
        lib.zpic_synchronize()
        log(f"batch: {min(done + 50, steps)} steps")
    t_batch = time.perf_counter() - t0
    e_batch = [member_energy(m) for m in members]
    b.close()
    for m in members:
        m.close()
    del members

    log("batch: sequential runs")
    t_seq = 0.0
    e_seq = []
    for s in seeds[::args.seq_stride]:
        m = scan.Member(lib, scan.member_params(s))
        if tiles and scan.size_class(m.params) in tiles:
            lib.zpic_sim_set_tiling(C.addressof(m.sim), *tiles[scan.size_class(m.params)])
        m.step(1)  # upload and first step outside the timed region, as for the batch
        lib.zpic_synchronize()
        t0 = time.perf_counter()
        m.step(steps - 1)
        lib.zpic_synchronize()
        t_seq += time.perf_counter() - t0
        e_seq.append(member_energy(m))
        m.close()
        if len(e_seq) % 64 == 0:
            log(f"batch: {len(e_seq)} sequential members")
    nseq = len(e_seq)
    # the sequential runs time steps 2..N; scale to N steps per member
    seq_rate = nseq / (t_seq * steps / (steps - 1))
    rel = [abs(a - c) / abs(c) for a, c in zip(e_batch[::args.seq_stride], e_seq) if c != 0]
    return {
        "metric": "batched sims/s",
        "value": len(seeds) / t_batch,
        "unit": "sims/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1e3 * t_batch / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded two-stream / Weibel members from the em2d injection",
        "config": {"workload": (f"parameter scan: {len(seeds)} two-stream/Weibel members, grids "
                                f"{{64,128,256}}^2 (independent x, y), 4-64 ppc per species, uz U[0.1,0.6], "
                                f"uth U[0.01,0.1], seeds 0..{len(seeds) - 1}, {steps} steps each"),
                   "members": len(seeds), "particles": npart},
        "particle_steps_per_s": npart * steps / t_batch,
        "sequential": {"value": seq_rate, "unit": "sims/s", "members_run": nseq,
                       "note": "same members, one after another with sim_iter on the GPU"},
        "speedup_vs_sequential": (len(seeds) / t_batch) / seq_rate,
        "tiles": {f"{k[0]}x{k[1]}": list(v) for k, v in tiles.items()} if tiles else "default 16x16, halo 3",
        "tuning_s": t_tune,
        "init": {"seconds": t_init, "where": "device" if device else f"host ({threads} threads)",
                 "setup_seconds": t_setup,
                 "note": "members' spec_new/sim_new, then zpic_batch_new (upload or device sort of every "
                         "member's particles); neither is in the timed region"},
        "energy_rel_diff_max": max(rel) if rel else None,
    }


def adjoint_workload(args) -> dict:
    """Differentiable 256^2 Weibel run (zpic_amd.diff defaults: 2 electron
    species of 4x4 ppc, 200 steps, checkpoints every 20, loss over steps
    150-200, 64x64 B_z seed): time of the checkpointed forward run and of
    the backward pass (re-run of every segment + reverse steps, up to the
    adjoints of the initial state in host memory).  The same run with the
    whole trajectory recorded by the forward run is reported beside it."""
    import dataclasses

    import numpy as np

    from zpic_amd import _lib
    from zpic_amd.diff import WeibelConfig, WeibelRun

    lib = _lib.load()
    if args.tile or args.tile_y or args.halo:
        lib.zpic_set_tiling(args.tile, args.tile_y or args.tile, args.halo)
    cfg = WeibelConfig()
    seed = 1e-3 * np.random.default_rng(0).standard_normal((cfg.ny // cfg.block, cfg.nx // cfg.block))
    # warm-up run (kernel loading, allocations)
    w = WeibelRun(WeibelConfig(nx=64, ny=64, steps=20, ckpt=10, loss_first=10, loss_last=20), 0.6, 0.1,
                  np.zeros((16, 16)))
    w.forward()
    w.gradients()
    w.close()

    def timed(c):
        log(f"adjoint: building the {c.nx}x{c.ny} run (record {c.record})")
        run = WeibelRun(c, 0.6, 0.1, seed)
        try:
            lib.zpic_synchronize()
            t0 = time.perf_counter()
            loss = run.forward()
            lib.zpic_synchronize()
            t_fwd = time.perf_counter() - t0
            log(f"adjoint: forward {t_fwd:.3f} s; backward")
            t0 = time.perf_counter()
            out = run.backward()
            t_bwd = time.perf_counter() - t0
            g_uz, g_uth, g_seed = run.reduce(*out)
            return {"npart": sum(run.np), "t_fwd": t_fwd, "t_bwd": t_bwd, "loss": loss,
                    "grad": {"uz": g_uz, "uth": g_uth, "seed_norm": float(np.linalg.norm(g_seed))}}
        finally:
            run.close()

    ck = timed(cfg)
    rec = timed(dataclasses.replace(cfg, record=True))
    return {
        "metric": "adjoint/fwd",
        "value": ck["t_bwd"] / ck["t_fwd"],
        "unit": "ratio of backward to forward time",
        "n_gpus": 1,
        "steps": cfg.steps,
        "warmup": 0,
        "ms_per_step": 1e3 * ck["t_bwd"] / cfg.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Weibel initial state, seeded 64x64 B_z perturbation (amplitude 1e-3)",
        "config": {"workload": (f"differentiable Weibel {cfg.nx}x{cfg.ny}, 2 electron species {cfg.ppc}x{cfg.ppc} "
                                f"ppc ({ck['npart']} particles), {cfg.steps} steps, checkpoints every {cfg.ckpt}, "
                                f"loss = sum of dt * B energy over steps {cfg.loss_first}-{cfg.loss_last}")},
        "forward_ms_per_step": 1e3 * ck["t_fwd"] / cfg.steps,
        "backward_ms_per_step": 1e3 * ck["t_bwd"] / cfg.steps,
        "loss": ck["loss"],
        "grad": ck["grad"],
        "recorded": {"value": rec["t_bwd"] / rec["t_fwd"],
                     "forward_ms_per_step": 1e3 * rec["t_fwd"] / cfg.steps,
                     "backward_ms_per_step": 1e3 * rec["t_bwd"] / cfg.steps,
                     "backward_over_checkpointed_forward": rec["t_bwd"] / ck["t_fwd"],
                     "loss": rec["loss"], "grad": rec["grad"]},
    }


def lwfa_workload(args) -> dict:
    """Laser wakefield with a moving window and open y boundaries: 4000 x
    256 cells, dx = 0.05, dy = 0.2, dt = 0.045; electrons (q/m = -1, 4 x 2
    ppc, cold) from x = 20 on; Gaussian laser a0 = 2, omega0 = 10, W0 = 4,
    FWHM 2, start 17, focus 20, axis 25.6, polarization 0; 4-pass
    compensated filter along x; 2000 steps (10 untimed)."""
    import ctypes as C

    from zpic_amd import _lib
    from zpic_amd import em2d

    lib = _lib.load()
    lib.zpic_set_push_waves(args.waves)
    nx, ny, dx, dy, dt = 4000, 256, 0.05, 0.2, 0.045
    log("lwfa: building 4000x256 with moving window and open y boundaries")
    lib.zpic_set_device_init(0 if args.host_init else 1)
    e = em2d.Species("electrons", -1.0, [4, 2], density=em2d.Density(type="step", start=20.0))
    sim = em2d.Simulation([nx, ny], [nx * dx, ny * dy], dt, species=[e], mov_window=True,
                          smooth=em2d.Smooth(xtype="compensated", xlevel=4), open_y=True)
    sim.add_laser(em2d.Laser(type="gaussian", start=17.0, fwhm=2.0, a0=2.0, omega0=10.0, W0=4.0, focus=20.0,
                             axis=25.6, polarization=0.0))
    lib.zpic_set_device_init(0)
    npart = int(e._s.np)
    warm, steps = 10, 1990
    for _ in range(warm):
        sim.iter()
    lib.zpic_synchronize()
    n0 = lib.spec_npush()
    t0 = time.perf_counter()
    for k in range(steps):
        sim.iter()
        if (k + 1) % 500 == 0:
            log(f"lwfa: queued {k + 1} steps")
    lib.zpic_synchronize()
    el = time.perf_counter() - t0
    pushes = lib.spec_npush() - n0
    sim.close()
    return {
        "metric": "particle-steps/s", "value": pushes / el, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": 1e3 * el / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: STEP plasma and Gaussian laser from the em2d initialisers",
        "config": {"workload": "em2d LWFA 4000x256, dx=0.05 dy=0.2 dt=0.045, moving window, open (Mur) y "
                               "boundaries, 4x2 ppc from x=20, Gaussian laser a0=2 omega0=10 W0=4 fwhm=2, "
                               "compensated 4-pass x filter",
                   "particles_at_start": npart},
        "pushes": int(pushes),
    }


def secondary_metrics(args) -> dict:
    """The other headline metrics on their own workloads, summarised: the
    batched scan (512 members; every 4th member also run sequentially for
    the sims/s of sequential runs), the adjoint/forward time ratio of the
    differentiable 256^2 run, and the LWFA particle-steps/s.  Each is the
    same measurement as its --workload run."""
    out = {}
    b_args = argparse.Namespace(**vars(args))
    b_args.members, b_args.batch_steps, b_args.seq_stride, b_args.tune = 512, 300, 4, False
    log("bench: secondary metrics: batched scan")
    b = batch_workload(b_args)
    out["batch"] = {"sims_per_s": b["value"], "sequential_sims_per_s": b["sequential"]["value"],
                    "speedup_vs_sequential": b["speedup_vs_sequential"], "members": b["config"]["members"],
                    "sequential_members_run": b["sequential"]["members_run"],
                    "particle_steps_per_s": b["particle_steps_per_s"],
                    "energy_rel_diff_max": b["energy_rel_diff_max"], "init": b["init"],
                    "workload": b["config"]["workload"]}
    log("bench: secondary metrics: adjoint")
    a = adjoint_workload(args)
    out["adjoint"] = {"adjoint_over_forward": a["value"], "forward_ms_per_step": a["forward_ms_per_step"],
                      "backward_ms_per_step": a["backward_ms_per_step"],
                      "recorded_adjoint_over_forward": a["recorded"]["value"], "workload": a["config"]["workload"]}
    log("bench: secondary metrics: LWFA")
    w = lwfa_workload(args)
    out["lwfa"] = {"particle_steps_per_s": w["value"], "ms_per_step": w["ms_per_step"],
                   "workload": w["config"]["workload"]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=990)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ppc", type=int, default=4)
    ap.add_argument("--n-sort", type=int, default=20)
    ap.add_argument("--tile", type=int, default=0, help="tile width in cells (0: library default)")
    ap.add_argument("--tile-y", type=int, default=0, help="tile height in cells (0: the tile width)")
    ap.add_argument("--halo", type=int, default=0, help="tile drift margin in cells (0: library default)")
    ap.add_argument("--waves", type=int, default=0, help="push variant: 0 (default), 2 or 8")
    ap.add_argument("--fixed", type=int, default=-1, help="fixed-point current deposit: 1 on, 0 off (-1: library default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-nx", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=40, help="CPU sample steps (40 span two tile sorts)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="0: every core of the process's affinity mask, at most 16 (one GPU's share)")
    ap.add_argument("--host-init", action="store_true",
                    help="initialise the particles on the host (default: on the device, zpic_set_device_init)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the batch, adjoint and LWFA measurements of the default line")
    ap.add_argument("--workload", choices=["weibel", "batch", "adjoint", "lwfa"], default="weibel",
                    help="weibel: the 2048^2 Weibel step (default); batch: the batched parameter scan; "
                         "adjoint: backward/forward time of the differentiable 256^2 Weibel run")
    ap.add_argument("--members", type=int, default=512, help="batch: members of the scan")
    ap.add_argument("--batch-steps", type=int, default=300, help="batch: steps per member")
    ap.add_argument("--seq-stride", type=int, default=1, help="batch: run every n-th member sequentially")
    ap.add_argument("--threads", type=int, default=0, help="batch: host threads for init (0: min(16, cores))")
    ap.add_argument("--tune", action="store_true", help="batch: autotune the tile per grid size class first")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES through the HIP runtime "
                         "(calibrates the FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_probe:
        return copy_probe(args.copy_probe)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the time step would shard by domain decomposition, not implemented here)", file=sys.stderr)
        return 2

    ensure_built()
    if args.workload == "batch":
        print(json.dumps(batch_workload(args)), flush=True)
        return 0
    if args.workload == "adjoint":
        print(json.dumps(adjoint_workload(args)), flush=True)
        return 0
    if args.workload == "lwfa":
        print(json.dumps(lwfa_workload(args)), flush=True)
        return 0
    import ctypes as C

    from zpic_amd import _lib
    from zpic_amd import em2d

    lib = _lib.load()
    if args.tile or args.tile_y or args.halo:
        lib.zpic_set_tiling(args.tile, args.tile_y or args.tile, args.halo)
    lib.zpic_set_push_waves(args.waves)
    if args.fixed >= 0:
        lib.zpic_set_fixed_deposit(args.fixed)
    nx = args.nx
    ppc = args.ppc
    dx, dt, uz, uth = 0.1, 0.07, 0.6, 0.1
    lib.zpic_set_device_init(0 if args.host_init else 1)
    log(f"bench: building Weibel {nx}x{nx}, 2 species x {ppc}x{ppc} ppc "
        f"({'host' if args.host_init else 'device'} initialisation)")
    lib.zpic_synchronize()
    t_init = time.time()
    sp = [em2d.Species("electrons", -1.0, [ppc, ppc], ufl=[0.0, 0.0, uz], uth=[uth] * 3, n_sort=args.n_sort),
          em2d.Species("electrons2", -1.0, [ppc, ppc], ufl=[0.0, 0.0, -uz], uth=[uth] * 3, n_sort=args.n_sort)]
    sim = em2d.Simulation([nx, nx], [nx * dx, nx * dx], dt, species=sp,
                          smooth=em2d.Smooth(xtype="binomial", ytype="binomial", xlevel=1, ylevel=1))
    npart = sum(s._s.np for s in sp)
    lib.zpic_synchronize()
    init_s = time.time() - t_init
    lib.zpic_set_device_init(0)
    log(f"bench: init {init_s:.2f} s, {npart} particles; {args.warmup} warm-up steps")
    for _ in range(args.warmup):
        sim.iter()
    lib.zpic_synchronize()

    lib.zpic_kernel_timing_reset()
    lib.zpic_kernel_timing(1)
    st0 = (C.c_longlong * 4)()
    lib.zpic_staged_stats(st0)
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.iter()
        if (k + 1) % 200 == 0:
            log(f"bench: queued {k + 1} steps")
    lib.zpic_synchronize()
    t1 = time.perf_counter()
    lib.zpic_kernel_timing(0)
    elapsed = t1 - t0
    launches = C.c_longlong(0)
    push_s = lib.zpic_kernel_time(0, C.byref(launches))
    kt = {}
    for k, name in enumerate(["push", "sort", "current", "field"]):
        n = C.c_longlong(0)
        tot = lib.zpic_kernel_time(k, C.byref(n))
        kt[name] = {"ms_per_launch": 1e3 * tot / max(n.value, 1), "launches": n.value}
    overflow = [lib.spec_overflow_count(s._ptr) for s in sp]
    st1 = (C.c_longlong * 4)()
    lib.zpic_staged_stats(st1)
    kt["sort"].update({"staged": st1[0] - st0[0], "full_after_staged": st1[1] - st0[1],
                       "left_without_staging_room": st1[2] - st0[2], "arrivals_without_free_slot": st1[3] - st0[3]})
    sim.close()

    value = npart * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    cells = nx * nx
    nspec = len(sp)
    # one push launch covers one species, or all of them when sim_iter fuses
    # the species (the default): 56 B per particle (read and write the seven
    # particle fields) + 36 B per cell (read E and B, write J) per launch
    spl = max(1, round(nspec * args.steps / max(launches.value, 1)))
    bytes_per_launch = 56.0 * (npart / nspec) * spl + 36.0 * cells
    launches_per_step = launches.value / args.steps
    avg_push = push_s / max(launches.value, 1)
    achieved = bytes_per_launch / avg_push / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "push_traffic.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            if tj.get("nx") == nx and tj.get("ppc") == ppc and tj.get("species_per_launch", 1) == spl:
                traffic = tj.get("bytes_per_launch")
        except Exception:
            traffic = None

    # whole step: both species' push bytes + current reset (12 B/cell), two
    # filter passes (24 B/cell each) and the fused Yee update (60 B/cell)
    step_bytes = launches_per_step * bytes_per_launch + (12.0 + 2 * 24.0 + 60.0) * cells
    step_achieved = step_bytes / (ms_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        nproc = args.cpu_procs or host_cores()["share"]
        log(f"bench: CPU baseline on {nproc} host cores")
        cpu = cpu_baseline(args.cpu_nx, ppc, args.cpu_steps, nproc)

    secondary = None
    if not args.no_secondary:
        secondary = secondary_metrics(args)

    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Weibel initial state from the em2d injection (MWC + Box-Muller)",
        "config": {"workload": (f"em2d Weibel {nx}x{nx} periodic, dx=dy=0.1, dt=0.07, 2 electron species "
                                f"u_z=+-0.6 u_th=0.1, {ppc}x{ppc} ppc each ({npart} particles), "
                                f"sort every {args.n_sort} steps, 1-pass binomial smoothing x and y"),
                   "nx": nx, "ppc": ppc, "particles": npart, "steps_total": args.warmup + args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("push_species_kernel (push + deposit, all species of the step in one launch)"
                                if spl > 1 else "push_tile_kernel (push + deposit, one species)")
                     + " + push_global_kernel",
                     "species_per_launch": spl,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": 1e3 * avg_push},
        "step_roofline": {"bound": "hbm", "achieved": step_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_achieved / HBM_PEAK_GBS, "bytes_per_step": step_bytes},
        "cpu_baseline": cpu,
        "vs_cpu": (value / cpu["value"]) if cpu else None,
        "vs_cpu_single_core": (value / cpu["single_core_value"]) if cpu else None,
        "kernels": kt,
        "overflow_pushes": overflow,
        "init": {"seconds": init_s, "where": "host" if args.host_init else "device",
                 "note": "spec_new of both species and sim_new, before the first step (not in the timed region)"},
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
