"""Differentiable restatement of the em2d time step in PyTorch (any float
dtype; FP64 for reference gradients).

TEST INFRASTRUCTURE ONLY: a checker for the GPU library's hand-written
adjoint (zpic_diff_*).  Nothing in the product imports it.  Gradients come
from torch autograd through this plain restatement, independent of the
GPU library's reverse kernels.

One step follows sim_iter (em2d/simulation.c:45-56) for periodic grids
without current filter or external fields:
  * field gather on the staggered mesh (em2d/particles.c:1029-1071),
  * relativistic Boris push (em2d/particles.c:1125-1228),
  * zigzag charge-conserving deposit split at cell edges
    (dep_current_zamb, em2d/particles.c:773-924),
  * periodic fold of the guard-cell current (em2d/current.c:142-183),
  * Yee update B half step, E step, B half step and guard copies
    (em2d/emf.c:500-562, 573-637, 688-716).
Branch decisions (interpolation stagger, edges crossed) are taken per
particle with torch.where, so the result is differentiable almost
everywhere.  Particle positions are (cell, offset) pairs; the offset is the
differentiable part.  This restatement is pinned against the C oracle
(float32 runs of both agree; tests/test_diff.py).
"""
from __future__ import annotations

import torch


def _wrap(i, n):
    return torch.remainder(i, n)


class Grid:
    def __init__(self, nx: int, ny: int, dx: float, dy: float, dt: float, dtype=torch.float64):
        self.nx, self.ny, self.dx, self.dy, self.dt = nx, ny, dx, dy, dt
        self.dtype = dtype
        W, H = nx + 3, ny + 3
        # periodic image of every buffer node (buffer index b <-> cell b - 1)
        bi = torch.arange(W)
        bj = torch.arange(H)
        self.img_i = _wrap(bi - 1, nx)
        self.img_j = _wrap(bj - 1, ny)
        self.img_flat = (self.img_j[:, None] * nx + self.img_i[None, :]).reshape(-1)

    # ---- grids: tensors (ny + 3, nx + 3, 3), [j + 1, i + 1] ----
    def zeros(self):
        return torch.zeros(self.ny + 3, self.nx + 3, 3, dtype=self.dtype)

    def fill_guards(self, F):
        """Guard nodes take their periodic image (emf_update_gc)."""
        inner = F[1:self.ny + 1, 1:self.nx + 1]
        return inner[self.img_j][:, self.img_i]

    def fold(self, raw):
        """Current fold: every node gets the sum over all buffer nodes with
        the same periodic image (current_update's guard add + copy)."""
        s = torch.zeros(self.ny * self.nx, 3, dtype=raw.dtype).index_add(0, self.img_flat, raw.reshape(-1, 3))
        return s.reshape(self.ny, self.nx, 3)[self.img_j][:, self.img_i]

    def yee(self, E, B, J):
        nx, ny, dt = self.nx, self.ny, self.dt
        dt2 = dt / 2.0
        dtx2, dty2, dtx, dty = dt2 / self.dx, dt2 / self.dy, dt / self.dx, dt / self.dy
        pad = torch.nn.functional.pad

        def bhalf(E, B):
            # B nodes i, j in [-1, n]: buffer [0, n + 2)
            Ex, Ey, Ez = E[..., 0], E[..., 1], E[..., 2]
            dbx = -dty2 * (Ez[1:ny + 3, 0:nx + 2] - Ez[0:ny + 2, 0:nx + 2])
            dby = dtx2 * (Ez[0:ny + 2, 1:nx + 3] - Ez[0:ny + 2, 0:nx + 2])
            dbz = (-dtx2 * (Ey[0:ny + 2, 1:nx + 3] - Ey[0:ny + 2, 0:nx + 2]) +
                   dty2 * (Ex[1:ny + 3, 0:nx + 2] - Ex[0:ny + 2, 0:nx + 2]))
            d = torch.stack([dbx, dby, dbz], dim=-1)
            return B + pad(d, (0, 0, 0, 1, 0, 1))

        def efull(E, B, J):
            # E nodes i, j in [0, n + 1]: buffer [1, n + 3)
            Bx, By, Bz = B[..., 0], B[..., 1], B[..., 2]
            dex = dty * (Bz[1:ny + 3, 1:nx + 3] - Bz[0:ny + 2, 1:nx + 3]) - dt * J[1:ny + 3, 1:nx + 3, 0]
            dey = -dtx * (Bz[1:ny + 3, 1:nx + 3] - Bz[1:ny + 3, 0:nx + 2]) - dt * J[1:ny + 3, 1:nx + 3, 1]
            dez = (dtx * (By[1:ny + 3, 1:nx + 3] - By[1:ny + 3, 0:nx + 2]) -
                   dty * (Bx[1:ny + 3, 1:nx + 3] - Bx[0:ny + 2, 1:nx + 3]) - dt * J[1:ny + 3, 1:nx + 3, 2])
            d = torch.stack([dex, dey, dez], dim=-1)
            return E + pad(d, (0, 0, 1, 0, 1, 0))

        B1 = bhalf(E, B)
        E1 = efull(E, B1, J)
        B2 = bhalf(E1, B1)
        return self.fill_guards(E1), self.fill_guards(B2)

    def b_energy(self, B):
        inner = B[1:self.ny + 1, 1:self.nx + 1]
        return 0.5 * self.dx * self.dy * (inner.to(torch.float64) ** 2).sum()


def _gather(F, i, j, wa, wb):
    """Bilinear interpolation of one component grid F[j, i] (buffer
    indices) from base node (i, j) with weights wa, wb."""
    f00 = F[j, i]
    f10 = F[j, i + 1]
    f01 = F[j + 1, i]
    f11 = F[j + 1, i + 1]
    return (f00 * (1 - wa) + f10 * wa) * (1 - wb) + (f01 * (1 - wa) + f11 * wa) * wb


def gather(E, B, ix, iy, x, y):
    """E and B at the particles (em2d/particles.c:1029-1071)."""
    i, j = ix + 1, iy + 1
    lo_x, lo_y = x < 0.5, y < 0.5
    ih = i - lo_x.long()
    jh = j - lo_y.long()
    wxh = torch.where(lo_x, x + 0.5, x - 0.5)
    wyh = torch.where(lo_y, y + 0.5, y - 0.5)
    ex = _gather(E[..., 0], ih, j, wxh, y)
    ey = _gather(E[..., 1], i, jh, x, wyh)
    ez = _gather(E[..., 2], i, j, x, y)
    bx = _gather(B[..., 0], i, jh, x, wyh)
    by = _gather(B[..., 1], ih, j, wxh, y)
    bz = _gather(B[..., 2], ih, jh, wxh, wyh)
    return torch.stack([ex, ey, ez], -1), torch.stack([bx, by, bz], -1)


def boris(u, e, b, tem):
    e = e * tem
    a = u + e
    g = torch.sqrt(1 + (a * a).sum(-1, keepdim=True))
    t = b * (tem / g)
    s2 = 2 / (1 + (t * t).sum(-1, keepdim=True))
    v = a + torch.cross(a, t, dim=-1)
    s = t * s2
    a = a + torch.cross(v, s, dim=-1)
    return a + e


def _safe(d, m):
    return torch.where(m, d, torch.ones_like(d))


def _ysplit(s, dj, m):
    """ysplit of the reference on the particles where m holds: returns the
    cut segment s' and the new segment t (fields as dict of tensors)."""
    edge = (dj == 1).to(s["y1"].dtype)
    djf = dj.to(s["y1"].dtype)
    f = (s["y1"] - edge) / _safe(s["dy"], m)
    xc = s["x0"] + s["dx"] * (1 - f)
    t = {"x0": xc, "x1": s["x1"], "y0": 1 - edge, "y1": s["y1"] - djf, "dx": s["dx"] * f, "dy": s["dy"] * f,
         "q": s["q"] * f, "ix": s["ix"], "iy": s["iy"] + dj}
    sp = dict(s)
    sp.update({"y1": edge, "dy": s["dy"] * (1 - f), "dx": s["dx"] * (1 - f), "x1": xc, "q": s["q"] * (1 - f)})
    return sp, t


def _sel(m, a, b):
    """segment a where m, else b"""
    out = {}
    for k in a:
        va, vb = a[k], b[k]
        if not torch.is_tensor(va):
            va = torch.full_like(vb, va)
        if not torch.is_tensor(vb):
            vb = torch.full_like(va, vb)
        if va.dtype != vb.dtype:
            vb = vb.to(va.dtype)
        out[k] = torch.where(m, va, vb)
    return out


def deposit(grid: Grid, ix, iy, x, y, mx, my, qvz, di, dj, qnx, qny):
    """Raw current of the zigzag deposit (em2d/particles.c:773-924) as a
    guard-inclusive grid."""
    dt_ = x.dtype
    s0 = {"x0": x, "y0": y, "dx": mx, "dy": my, "x1": x + mx, "y1": y + my, "q": qvz * 0.5, "ix": ix, "iy": iy}
    xs = di != 0
    ys = dj != 0
    edge = (di == 1).to(dt_)
    fx = (x + mx - edge) / _safe(mx, xs)
    yc = y + my * (1 - fx)
    s1x = {"x0": 1 - edge, "x1": (x + mx) - di.to(dt_), "dx": mx * fx, "dy": my * fx, "y0": yc, "y1": s0["y1"],
           "q": s0["q"] * fx, "ix": ix + di, "iy": iy}
    s0x = dict(s0)
    s0x.update({"x1": edge, "dx": mx * (1 - fx), "dy": my * (1 - fx), "y1": yc, "q": s0["q"] * (1 - fx)})
    s0 = _sel(xs, s0x, s0)
    s1 = _sel(xs, s1x, s0)  # s1 only meaningful where xs or ys
    # y cuts
    case1 = ys & ~xs
    out_y = (s0["y1"] < 0) | (s0["y1"] >= 1)
    case2 = ys & xs & out_y
    case3 = ys & xs & ~out_y
    a_s, a_t = _ysplit(s0, dj, case1 | case2)  # cut of s0
    b_s, b_t = _ysplit(s1, dj, case3)  # cut of s1
    djf = dj.to(dt_)
    s1_shift = dict(s1)
    s1_shift.update({"y0": s1["y0"] - djf, "y1": s1["y1"] - djf, "iy": s1["iy"] + dj})
    new0 = _sel(case1 | case2, a_s, s0)
    new1 = _sel(case1, a_t, _sel(case2, s1_shift, _sel(case3, b_s, s1)))
    new2 = _sel(case2, a_t, b_t)
    v1 = xs | ys
    v2 = xs & ys
    raw = grid.zeros().reshape(-1)
    W = grid.nx + 3
    for seg, valid in ((new0, None), (new1, v1), (new2, v2)):
        ax0, bx0, ax1, bx1 = 1 - seg["x0"], seg["x0"], 1 - seg["x1"], seg["x1"]
        ay0, by0, ay1, by1 = 1 - seg["y0"], seg["y0"], 1 - seg["y1"], seg["y1"]
        jx = qnx * seg["dx"]
        jy = qny * seg["dy"]
        wy0, wy1 = 0.5 * (ay0 + ay1), 0.5 * (by0 + by1)
        wx0, wx1 = 0.5 * (ax0 + ax1), 0.5 * (bx0 + bx1)
        q = seg["q"]
        terms = [
            (0, 0, 0, jx * wy0), (0, 1, 0, jx * wy1),
            (0, 0, 1, jy * wx0), (1, 0, 1, jy * wx1),
            (0, 0, 2, q * (ax0 * ay0 + ax1 * ay1 + (ax0 * ay1 - ax1 * ay0) / 2)),
            (1, 0, 2, q * (bx0 * ay0 + bx1 * ay1 + (bx0 * ay1 - bx1 * ay0) / 2)),
            (0, 1, 2, q * (ax0 * by0 + ax1 * by1 + (ax0 * by1 - ax1 * by0) / 2)),
            (1, 1, 2, q * (bx0 * by0 + bx1 * by1 + (bx0 * by1 - bx1 * by0) / 2)),
        ]
        for a, b, c, val in terms:
            if valid is not None:
                val = torch.where(valid, val, torch.zeros_like(val))
            node = (seg["iy"] + 1 + b) * W + (seg["ix"] + 1 + a)
            raw = raw.index_add(0, 3 * node + c, val)
    return raw.reshape(grid.ny + 3, grid.nx + 3, 3)


class Species:
    def __init__(self, q: float, m_q: float, ix, iy, x, y, u):
        self.q, self.m_q = q, m_q
        self.ix, self.iy = ix.long(), iy.long()
        self.x, self.y, self.u = x, y, u


def step(grid: Grid, species: list, E, B):
    """One sim_iter; species are updated in place (new tensors), returns
    the new (E, B)."""
    dt = grid.dt
    raw = grid.zeros()
    for s in species:
        tem = 0.5 * dt / s.m_q
        e, b = gather(E, B, s.ix, s.iy, s.x, s.y)
        u = boris(s.u, e, b, tem)
        rg = 1 / torch.sqrt(1 + (u * u).sum(-1))
        mx = (dt / grid.dx) * rg * u[:, 0]
        my = (dt / grid.dy) * rg * u[:, 1]
        x1, y1 = s.x + mx, s.y + my
        di = (x1 >= 1).long() - (x1 < 0).long()
        dj = (y1 >= 1).long() - (y1 < 0).long()
        qvz = s.q * u[:, 2] * rg
        raw = raw + deposit(grid, s.ix, s.iy, s.x, s.y, mx, my, qvz, di, dj, s.q * grid.dx / dt, s.q * grid.dy / dt)
        s.x = x1 - di.to(x1.dtype)
        s.y = y1 - dj.to(y1.dtype)
        s.ix = _wrap(s.ix + di, grid.nx)
        s.iy = _wrap(s.iy + dj, grid.ny)
        s.u = u
    J = grid.fold(raw)
    return grid.yee(E, B, J)


def run_loss(grid: Grid, species: list, E, B, steps: int, loss_first: int, loss_last: int):
    """The loss of zpic_diff_forward: sum over n in [loss_first, loss_last]
    of dt * (B field energy after n steps)."""
    loss = torch.zeros((), dtype=torch.float64)
    if loss_first <= 0 <= loss_last:
        loss = loss + grid.dt * grid.b_energy(B)
    for n in range(steps):
        E, B = step(grid, species, E, B)
        if loss_first <= n + 1 <= loss_last:
            loss = loss + grid.dt * grid.b_energy(B)
    return loss, E, B
