"""Seeded synthetic inputs of the benchmark shapes (numpy Generator streams).

blobs_f32     Gaussian blobs, L2-normalised rows (f32 configs)
blobs_i8      normalised blob vectors scaled by 127, rounded, clipped (int8 config)
bits_u8       Bernoulli prototypes with per-bit flips, packed 8 bits per byte
"""
from __future__ import annotations

import numpy as np


def _sizes(n: int, centers: int, zipf: float | None) -> np.ndarray:
    if zipf is None:
        w = np.ones(centers)
    else:
        w = np.arange(1, centers + 1, dtype=np.float64) ** (-zipf)
    sizes = np.floor(w / w.sum() * n).astype(np.int64)
    sizes[: n - int(sizes.sum())] += 1
    return sizes


def blobs_f32(n: int, d: int, centers: int, std: float, seed: int, zipf: float | None = None,
              normalize: bool = True) -> np.ndarray:
    """centers ~ N(0, I), row = center + std * N(0, I), rows in shuffled blob order."""
    g = np.random.default_rng(seed)
    c = g.standard_normal((centers, d), dtype=np.float32)
    lab = np.repeat(np.arange(centers), _sizes(n, centers, zipf))
    g.shuffle(lab)
    x = np.empty((n, d), np.float32)
    step = 1 << 16
    for s in range(0, n, step):
        e = min(n, s + step)
        x[s:e] = c[lab[s:e]] + np.float32(std) * g.standard_normal((e - s, d), dtype=np.float32)
    if normalize:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x


def blobs_i8(n: int, d: int, centers: int, std: float, seed: int) -> np.ndarray:
    x = blobs_f32(n, d, centers, std, seed)
    return np.clip(np.rint(x * 127.0), -127, 127).astype(np.int8)


def bits_u8(n: int, bits: int, prototypes: int, p_one: float, p_flip: float, seed: int) -> np.ndarray:
    g = np.random.default_rng(seed)
    protos = g.random((prototypes, bits)) < p_one
    out = np.empty((n, bits // 8), np.uint8)
    step = 1 << 16
    for s in range(0, n, step):
        e = min(n, s + step)
        b = protos[g.integers(0, prototypes, e - s)] ^ (g.random((e - s, bits)) < p_flip)
        out[s:e] = np.packbits(b, axis=1)
    return out
