"""Differentiable em2d runs behind torch.autograd.Function.

No counterpart in the reference.  `WeibelLoss.apply(uz, uth, bz_seed)`
returns the time-integrated magnetic field energy of a two-stream Weibel
run on the GPU, and its backward pass returns the gradients with respect
to the drift uz, the thermal spread uth and a coarse seed of the initial
B_z, computed by the library's adjoint kernels (zpic_diff_*).

The run (WeibelConfig): nx x ny periodic grid, two electron species
(q/m = -1) with p x p particles per cell injected by the em2d generator
(spec_new, set_rand_seed); their momenta are u = (0, 0, +-uz) + uth * G,
where G is the thermal deviate of the injection with unit spread and zero
drift (spec_set_u, em2d/particles.c:85-149, with its per-cell mean
removed), so the momenta depend linearly on (uz, uth).  The initial B_z on
every node, guard cells included (emf_init_fld, em2d/emf.c:946-978), is the
coarse seed repeated over blocks of `block` x `block` nodes:
B_z(i, j) = seed[(j mod ny) // block, (i mod nx) // block].
The loss is  sum over n in [loss_first, loss_last] of dt * (E_B3 + E_B4 +
E_B5)(n), the magnetic terms of emf_get_energy after n steps, in FP64.

PyTorch is used only for the autograd interface: the simulation, the
checkpointed forward run and the reverse kernels are the library's.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import abi as A

_libc = C.CDLL(None)
_libc.calloc.restype = C.c_void_p
_libc.calloc.argtypes = [C.c_size_t, C.c_size_t]


@dataclass(frozen=True)
class WeibelConfig:
    nx: int = 256
    ny: int = 256
    dx: float = 0.1
    dt: float = 0.07
    ppc: int = 4
    steps: int = 200
    ckpt: int = 20
    loss_first: int = 150
    loss_last: int = 200
    block: int = 4
    seed: tuple = (12345, 67890)
    # keep the state of every step during the forward run (memory of the
    # whole trajectory) instead of running each segment again backwards
    record: bool = False


def _declare(lib):
    if getattr(lib, "_diff_declared", False):
        return
    P = C.POINTER
    lib.zpic_diff_new.restype = C.c_void_p
    lib.zpic_diff_new.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.zpic_diff_forward.restype = C.c_double
    lib.zpic_diff_forward.argtypes = [C.c_void_p]
    lib.zpic_diff_backward.restype = None
    lib.zpic_diff_backward.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), P(P(C.c_float))]
    lib.zpic_diff_delete.restype = None
    lib.zpic_diff_delete.argtypes = [C.c_void_p]
    lib.zpic_diff_record_all.restype = None
    lib.zpic_diff_record_all.argtypes = [C.c_void_p, C.c_int]
    lib._diff_declared = True


class WeibelRun:
    """One differentiable run on the GPU: build, forward, backward, free."""

    def __init__(self, cfg: WeibelConfig, uz: float, uth: float, bz_seed: np.ndarray):
        lib = _lib.load()
        _declare(lib)
        self.lib, self.cfg = lib, cfg
        abi = A.make_abi(True)
        nx = (C.c_int * 2)(cfg.nx, cfg.ny)
        box = (C.c_float * 2)(cfg.nx * cfg.dx, cfg.ny * cfg.dx)
        dt = C.c_float(cfg.dt)
        lib.zpic_rand_clear()
        lib.set_rand_seed(cfg.seed[0], cfg.seed[1])
        raw = _libc.calloc(2, C.sizeof(abi.t_species))
        self.species = C.cast(raw, C.POINTER(abi.t_species))
        self.G = []
        self.sign = (1.0, -1.0)
        for k, name in enumerate(("electrons", "electrons2")):
            lib.spec_new(C.byref(self.species[k]), name.encode(), C.c_float(-1.0),
                         (C.c_int * 2)(cfg.ppc, cfg.ppc), (C.c_float * 3)(0.0, 0.0, 0.0),
                         (C.c_float * 3)(1.0, 1.0, 1.0), nx, box, dt, None)
            p = A.particles_view(self.species[k])
            G = np.stack([p["ux"], p["uy"], p["uz"]], axis=1).astype(np.float32)
            self.G.append(G)
            p["ux"] = np.float32(uth) * G[:, 0]
            p["uy"] = np.float32(uth) * G[:, 1]
            p["uz"] = np.float32(self.sign[k] * uz) + np.float32(uth) * G[:, 2]
        self.sim = abi.t_simulation()
        lib.sim_new(C.byref(self.sim), nx, box, dt, C.c_float(0.0), 0, self.species, 2)
        # guard-inclusive B grid: node (i, j), i, j in [-1, n + 1], at [j + 1, i + 1]
        full = np.ctypeslib.as_array(C.cast(self.sim.emf.B_buf, C.POINTER(C.c_float)),
                                     shape=(cfg.ny + 3, cfg.nx + 3, 3))
        full[..., 2] = self.upsample(np.asarray(bz_seed, dtype=np.float32))
        lib.emf_mark_host_dirty(C.byref(self.sim.emf))
        for k in range(2):
            lib.spec_mark_host_dirty(C.byref(self.species[k]))
        self.np = [int(self.species[k].np) for k in range(2)]
        self.q = [float(self.species[k].q) for k in range(2)]
        self.initial = [A.particles_view(self.species[k]).copy() for k in range(2)]
        self.d = lib.zpic_diff_new(C.byref(self.sim), cfg.steps, cfg.ckpt, cfg.loss_first, cfg.loss_last)
        if cfg.record:
            lib.zpic_diff_record_all(self.d, 1)

    def block_index(self):
        cfg = self.cfg
        i = np.remainder(np.arange(cfg.nx + 3) - 1, cfg.nx) // cfg.block
        j = np.remainder(np.arange(cfg.ny + 3) - 1, cfg.ny) // cfg.block
        return j[:, None], i[None, :]

    def upsample(self, seed: np.ndarray) -> np.ndarray:
        jj, ii = self.block_index()
        return seed[jj, ii]

    def forward(self) -> float:
        return float(self.lib.zpic_diff_forward(self.d))

    def backward(self):
        """dL/d(initial B grid) and the per-particle adjoints."""
        cfg = self.cfg
        nodes = (cfg.nx + 3) * (cfg.ny + 3)
        dB = np.zeros(3 * nodes, dtype=np.float32)
        dE = np.zeros(3 * nodes, dtype=np.float32)
        parts = [np.zeros(5 * n, dtype=np.float32) for n in self.np]
        FP = C.POINTER(C.c_float)
        arr = (FP * 2)(*[p.ctypes.data_as(FP) for p in parts])
        self.lib.zpic_diff_backward(self.d, dE.ctypes.data_as(FP), dB.ctypes.data_as(FP), arr)
        return (dE.reshape(cfg.ny + 3, cfg.nx + 3, 3), dB.reshape(cfg.ny + 3, cfg.nx + 3, 3),
                [p.reshape(n, 5) for p, n in zip(parts, self.np)])

    def gradients(self):
        """(dL/duz, dL/duth, dL/dseed) from the adjoint of the initial state."""
        return self.reduce(*self.backward())

    def reduce(self, dE, dB, parts):
        """(dL/duz, dL/duth, dL/dseed) from the output of backward()."""
        g_uz = sum(self.sign[k] * float(parts[k][:, 4].astype(np.float64).sum()) for k in range(2))
        g_uth = sum(float((parts[k][:, 2:5].astype(np.float64) * self.G[k]).sum()) for k in range(2))
        cfg = self.cfg
        jj, ii = self.block_index()
        g_seed = np.zeros((cfg.ny // cfg.block, cfg.nx // cfg.block), dtype=np.float64)
        np.add.at(g_seed, (np.broadcast_to(jj, dB.shape[:2]), np.broadcast_to(ii, dB.shape[:2])),
                  dB[..., 2].astype(np.float64))
        return g_uz, g_uth, g_seed

    def close(self):
        if self.d:
            self.lib.zpic_diff_delete(self.d)
            self.d = None
        if self.sim is not None:
            self.lib.sim_delete(C.byref(self.sim))
            self.sim = None


class WeibelLoss(torch.autograd.Function):
    """loss = WeibelLoss.apply(uz, uth, bz_seed, cfg): a float64 scalar."""

    @staticmethod
    def forward(ctx, uz: torch.Tensor, uth: torch.Tensor, bz_seed: torch.Tensor, cfg: WeibelConfig = WeibelConfig()):
        run = WeibelRun(cfg, float(uz), float(uth), bz_seed.detach().cpu().numpy())
        loss = run.forward()
        ctx.run = run
        ctx.seed_meta = (bz_seed.dtype, bz_seed.device)
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        run = ctx.run
        try:
            g_uz, g_uth, g_seed = run.gradients()
        finally:
            run.close()
        dtype, device = ctx.seed_meta
        g = float(grad)
        return (torch.tensor(g * g_uz, dtype=torch.float64), torch.tensor(g * g_uth, dtype=torch.float64),
                torch.from_numpy(g * g_seed).to(dtype=dtype, device=device), None)


def weibel_loss(uz, uth, bz_seed, cfg: WeibelConfig = WeibelConfig()):
    return WeibelLoss.apply(uz, uth, bz_seed, cfg)
