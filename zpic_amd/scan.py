"""Parameter-scan decks: many independent two-stream / Weibel simulations.

Member `seed` of a scan is two counter-streaming electron species
(q/m = -1, drifts +uz and -uz, thermal spread uth in x, y, z; the deck of
em2d/input/weibel.c:10-42 with two electron populations) on an nx x ny
periodic grid, dx = dy = 0.1, dt = 0.07, p x p particles per cell each,
tile sort every 20 steps, no current filter.  The member's parameters are
drawn from numpy.random.default_rng(seed):

    nx, ny  uniform over {64, 128, 256} (independently)
    p       uniform over {2, ..., 8}         (4 to 64 particles per cell)
    uz      uniform in [0.1, 0.6]
    uth     uniform in [0.01, 0.1]

and its particles from the em2d generator seeded with
set_rand_seed(12345 + seed, 67890 + seed) (em2d/random.c:16-17 defaults
offset by the seed).  Members are built directly through the em2d C API,
on several host threads at once: each worker thread draws from a
generator state of its own (zpic_rand_thread_private), so each member's
initial state is the one a single-threaded run would make.
"""
from __future__ import annotations

import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from . import abi as A

GRIDS = (64, 128, 256)
DX = 0.1
DT = 0.07
N_SORT = 20

_libc = C.CDLL(None)
_libc.calloc.restype = C.c_void_p
_libc.calloc.argtypes = [C.c_size_t, C.c_size_t]


def member_params(seed: int) -> dict:
    rng = np.random.default_rng(seed)
    nx = int(rng.choice(GRIDS))
    ny = int(rng.choice(GRIDS))
    p = int(rng.integers(2, 9))
    uz = float(rng.uniform(0.1, 0.6))
    uth = float(rng.uniform(0.01, 0.1))
    return {"seed": seed, "nx": nx, "ny": ny, "ppc": p, "uz": uz, "uth": uth}


class Member:
    """One simulation of a scan, owned through the em2d C structs."""

    def __init__(self, lib: C.CDLL, params: dict):
        abi = A.make_abi(True)
        self.lib = lib
        self.params = params
        nx = (C.c_int * 2)(params["nx"], params["ny"])
        box = (C.c_float * 2)(params["nx"] * DX, params["ny"] * DX)
        dt = C.c_float(DT)
        lib.zpic_rand_clear()
        lib.set_rand_seed(12345 + params["seed"], 67890 + params["seed"])
        raw = _libc.calloc(2, C.sizeof(abi.t_species))
        self.species = C.cast(raw, C.POINTER(abi.t_species))
        p, uz, uth = params["ppc"], params["uz"], params["uth"]
        for k, (name, drift) in enumerate((("electrons", uz), ("electrons2", -uz))):
            lib.spec_new(C.byref(self.species[k]), name.encode(), C.c_float(-1.0), (C.c_int * 2)(p, p),
                         (C.c_float * 3)(0.0, 0.0, drift), (C.c_float * 3)(uth, uth, uth), nx, box, dt, None)
            self.species[k].n_sort = N_SORT
        self.sim = abi.t_simulation()
        lib.sim_new(C.byref(self.sim), nx, box, dt, C.c_float(0.0), 0, self.species, 2)

    @property
    def particles(self) -> int:
        return int(self.species[0].np + self.species[1].np)

    def step(self, n: int = 1):
        for _ in range(n):
            self.lib.sim_iter(C.byref(self.sim))

    def close(self):
        if self.sim is not None:
            self.lib.sim_delete(C.byref(self.sim))
            self.sim = None


def build_members(seeds, threads: int = 16, tiles: dict | None = None, device: bool = False) -> list[Member]:
    """Members for the given seeds, initialised on `threads` host threads,
    or with `device` on the GPU (zpic_set_device_init: the same particles,
    made member after member on this thread, which then stay on the device
    until the batch sorts them); `tiles` maps a size class (nx, ny) to the
    tile (tx, ty, halo) to use."""
    lib = _lib.load()
    seeds = list(seeds)

    def make(s):
        if not device:
            lib.zpic_rand_thread_private(1)
        m = Member(lib, member_params(s))
        if tiles:
            t = tiles.get(size_class(m.params))
            if t:
                lib.zpic_sim_set_tiling(C.addressof(m.sim), *t)
        return m

    if device:
        prev = lib.zpic_device_init()
        lib.zpic_set_device_init(1)
        try:
            return [make(s) for s in seeds]
        finally:
            lib.zpic_set_device_init(prev)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(make, seeds))


# ---- tile autotuning per size class ----

# A batch launches its push with the largest LDS window of its members, so
# the candidates keep the window within the default's (16 + 2*3 + 3 = 25).
TILE_CANDIDATES = ((16, 16, 3), (16, 16, 2), (16, 8, 3), (8, 16, 3), (8, 8, 3), (8, 8, 2))


def size_class(params: dict) -> tuple:
    return (params["nx"], params["ny"])


def tune_tiles(classes, members: int = 8, steps: int = 10, candidates=TILE_CANDIDATES) -> dict:
    """The fastest tile of `candidates` for each size class (nx, ny),
    timed on a batch of `members` mid-range members of that class (p = 5,
    uz = 0.35, uth = 0.05; generator seeds outside the scan's) for `steps`
    steps after one warm-up step."""
    import time

    lib = _lib.load()
    best = {}
    for nx, ny in sorted(set(classes)):
        times = {}
        for cand in candidates:
            ms = []
            for k in range(members):
                m = Member(lib, {"seed": 100000 + k, "nx": nx, "ny": ny, "ppc": 5, "uz": 0.35, "uth": 0.05})
                lib.zpic_sim_set_tiling(C.addressof(m.sim), *cand)
                ms.append(m)
            b = Batch(ms)
            b.run(1)
            lib.zpic_synchronize()
            t0 = time.perf_counter()
            b.run(steps)
            lib.zpic_synchronize()
            times[cand] = time.perf_counter() - t0
            b.close()
            for m in ms:
                m.close()
        best[(nx, ny)] = min(times, key=times.get)
    return best


class Batch:
    """zpic_batch over scan members (or any em2d simulations)."""

    def __init__(self, members):
        lib = _lib.load()
        self.lib = lib
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[C.addressof(m.sim) for m in self.members])
        self.b = lib.zpic_batch_new(arr, len(self.members))

    def run(self, steps: int):
        self.lib.zpic_batch_run(self.b, int(steps))

    def close(self):
        if self.b:
            self.lib.zpic_batch_delete(self.b)
            self.b = None
