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


# ---- device-side generators (torch on the GPU) for the million-row shapes ----

def blobs_f32_device(n: int, d: int, centers: int, std: float, seed: int, zipf: float | None = None):
    """blobs_f32 drawn on the GPU (torch), L2-normalised, float32 [n, d] on cuda:0."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    c = torch.randn((centers, d), generator=g, device="cuda")
    sizes = torch.from_numpy(_sizes(n, centers, zipf)).cuda()
    lab = torch.repeat_interleave(torch.arange(centers, device="cuda"), sizes)
    lab = lab[torch.randperm(n, generator=g, device="cuda")]
    x = torch.empty((n, d), dtype=torch.float32, device="cuda")
    step = 1 << 17
    for s in range(0, n, step):
        e = min(n, s + step)
        x[s:e] = c[lab[s:e]] + std * torch.randn((e - s, d), generator=g, device="cuda")
    return torch.nn.functional.normalize(x, dim=1)


def blobs_i8_device(n: int, d: int, centers: int, std: float, seed: int):
    """Normalised blob vectors x 127, rounded and clipped: int8 [n, d] on cuda:0."""
    import torch

    x = blobs_f32_device(n, d, centers, std, seed)
    return torch.clamp(torch.round(x * 127.0), -127, 127).to(torch.int8)


def bits_u8_device(n: int, bits: int, prototypes: int, p_one: float, p_flip: float, seed: int):
    """Bernoulli(p_one) prototypes, each row a random prototype with every bit
    flipped with probability p_flip, packed MSB-first (numpy.packbits order):
    uint8 [n, bits / 8] on cuda:0."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    protos = torch.rand((prototypes, bits), generator=g, device="cuda") < p_one
    weights = (2 ** torch.arange(7, -1, -1, device="cuda")).to(torch.int32)
    out = torch.empty((n, bits // 8), dtype=torch.uint8, device="cuda")
    step = 1 << 16
    for s in range(0, n, step):
        e = min(n, s + step)
        pick = torch.randint(0, prototypes, (e - s,), generator=g, device="cuda")
        b = protos[pick] ^ (torch.rand((e - s, bits), generator=g, device="cuda") < p_flip)
        out[s:e] = (b.view(e - s, bits // 8, 8).to(torch.int32) * weights).sum(dim=2).to(torch.uint8)
    return out
