"""Seeded random linear sketches of half-spectrum fields.

A sketch of a complex field u over S storage slots is s_j = sum_i w_j(i) u(i)
with real weights w_j(i) in [-1, 1) from a counter-based hash of (seed, j, i),
so the writer of a fixture and its checker regenerate the same weights without
storing them. For an error field e, E|sum_i w_j(i) e(i)|^2 = |e|^2 / 3, so the
relative error of K such sketches estimates the relative L2 error of the whole
field (phase, sign and index errors included) to within a small factor, where
low-mode coefficients or a shell spectrum see only part of it.

The weights are w = (h >> 40) * 2^-23 - 1 with h = SplitMix64(i + S*j + 2^40 *
seed) (Steele, Lea & Flood 2014): 24-bit values, exact in float32.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = x + _GOLDEN
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        return x ^ (x >> np.uint64(31))


def weights(slots: int, count: int, seed: int = 7) -> np.ndarray:
    """[count, slots] float32 weight matrix (exact values)."""
    idx = np.arange(slots, dtype=np.uint64)
    out = np.empty((count, slots), np.float32)
    base = np.uint64(seed) << np.uint64(40)
    for j in range(count):
        with np.errstate(over="ignore"):
            h = splitmix64(idx + np.uint64(slots) * np.uint64(j) + base)
        out[j] = (h >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -23) - np.float32(1.0)
    return out


def sketch(w: np.ndarray, field: np.ndarray, chunk: int = 1 << 20) -> np.ndarray:
    """Complex sketches [count] of a complex field (any shape, C order)."""
    f = np.ascontiguousarray(field, np.complex128).reshape(-1).view(np.float64).reshape(-1, 2)
    if f.shape[0] != w.shape[1]:
        raise ValueError(f"field has {f.shape[0]} slots, weights {w.shape[1]}")
    acc = np.zeros((w.shape[0], 2))
    for i in range(0, f.shape[0], chunk):
        acc += w[:, i:i + chunk].astype(np.float64) @ f[i:i + chunk]
    return acc[:, 0] + 1j * acc[:, 1]


def low_block(half: np.ndarray, k: int) -> np.ndarray:
    """Modes |mx|, |my|, |mz| <= k of a half-spectrum [kx][kz][ky] field,
    as [kx 0..k][kz -k..k][ky -k..k]."""
    n = half.shape[1]
    idx = [m % n for m in range(-k, k + 1)]
    return np.ascontiguousarray(half[: k + 1][:, idx][:, :, idx])
