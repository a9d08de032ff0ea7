"""The push's fixed-point current windows, on the CPU.

Each contribution v (FP32) to a window node becomes the bit pattern of the
FP64 sum v * scale + 1.5 * 2^52, rounded once (zk_fx_raw in
zpic_amd/csrc/kernels/common.h: the product is exact, so one rounding, as
in the device's fused multiply-add); the window adds these patterns with
64-bit integer atomics and reads the sum back from its low 51 bits,
sign-extended (zk_fx_read).  Checked here against the sum of the rounded
integers computed directly, at sizes up to the tile limit the push
enforces (ZK_FX_TILE_MAX particles, each adding at most Q to a node).
"""
from __future__ import annotations

import numpy as np

MAGIC = 6755399441055744.0  # 1.5 * 2^52
TILE_MAX = (1 << 18) - 16   # ZK_FX_TILE_MAX in zpic_amd/csrc/kernels/zk.h


def raw(v: np.ndarray, scale: float) -> np.ndarray:
    r = v.astype(np.float64) * scale + MAGIC  # exact product, one rounding
    return r.view(np.uint64)


def read(s: np.uint64) -> int:
    return int((np.uint64(s) << np.uint64(13)).view(np.int64) >> np.int64(13))


def rounded(v: np.ndarray, scale: float) -> np.ndarray:
    """round-half-even of v * scale, as exact integers"""
    return np.rint(v.astype(np.float64) * scale).astype(np.int64)


def wrapsum(x: np.ndarray) -> np.uint64:
    with np.errstate(over="ignore"):
        return np.add.reduce(x, dtype=np.uint64)


def test_sum_of_patterns_reads_back_the_integer_sum():
    rng = np.random.default_rng(7)
    q = 2.0 ** -4  # charge of a 16-ppc species
    scale = 2.0 ** 32 / q
    for n in (1, 2, 3, 1000, 65536, TILE_MAX):
        # contributions of either sign, at most q (one per particle)
        v = (rng.uniform(-1.0, 1.0, n) * q).astype(np.float32)
        v[: n // 7] = np.float32(q)  # some at the bound
        k = rounded(v, scale)
        assert read(wrapsum(raw(v, scale))) == int(k.sum(dtype=np.int64))


def test_sum_at_the_range_bound():
    """Every particle of the largest tile adding the full Q with one sign:
    |sum| just below 2^50, still read back exactly; twice as many wrap."""
    q = 2.0 ** -3
    scale = 2.0 ** 32 / q
    for sign in (1.0, -1.0):
        v = np.full(TILE_MAX, sign * q, np.float32)
        s = read(wrapsum(raw(v, scale)))
        assert s == int(sign * TILE_MAX * 2 ** 32)
        assert abs(s) < 2 ** 50
        v2 = np.full(2 * TILE_MAX + 64, sign * q, np.float32)
        assert read(wrapsum(raw(v2, scale))) != int(sign * len(v2) * 2 ** 32)


def test_read_back_needs_the_sign_extension():
    """The raw sum itself is not the integer sum: N patterns carry
    N * 2151 * 2^51 in the bits above the 51 that are read."""
    q = 1.0
    scale = 2.0 ** 32
    v = np.array([0.25, -0.5, 0.125], np.float32)
    s = wrapsum(raw(v, scale))
    exact = int(rounded(v, scale).sum())
    assert read(s) == exact
    assert int(s.view(np.int64)) != exact
    # the bits above the low 51 are N * 2151 mod 2^13
    assert (int(s) >> 51) == (3 * 2151 + (exact >> 51)) % (1 << 13)


def test_rounding_is_to_nearest_even_at_the_scale():
    scale = 2.0 ** 32
    v = np.array([2.0 ** -33, 3 * 2.0 ** -33, -(2.0 ** -33), -3 * 2.0 ** -33, 2.0 ** -34], np.float32)
    # v * scale = 0.5, 1.5, -0.5, -1.5, 0.25: ties to even, then nearest
    got = [read(x) for x in raw(v, scale)]
    assert got == [0, 2, 0, -2, 0]
