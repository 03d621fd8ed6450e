"""Seeded synthetic inputs for the solver workloads.

* single solve: 64^3 (or any box), log-normal (mu=0, sigma) diagonal
  conductivity per cell, a +1/-1 point-current dipole at two seeded interior cells;
* batched uneven solves: edges drawn from {32, 48, 64, 96, 128}, per-instance
  sigma in [0.5, 2], 1-4 dipoles;
* 2D periodic flow: a divergence-free start from random Fourier modes.

``host_fields`` draws an instance from ``numpy.random.default_rng(seed)``
(PCG64) in a fixed order, so any instance of a workload can be regenerated
exactly on the host; ``device_fields`` is a faster GPU draw from a seeded
``torch.Generator`` for inputs that need no host twin.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from typing import List, Tuple

import numpy as np

EDGES = (32, 48, 64, 96, 128)


def batched_specs(n: int, seed: int, edges=EDGES) -> List[Tuple[Tuple[int, int, int], float, int, int]]:
    """Per-instance ((D,H,W), sigma_ln, n_dipoles, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        e = int(rng.choice(edges))
        s = float(rng.uniform(0.5, 2.0))
        k = int(rng.integers(1, 5))
        out.append(((e, e, e), s, k, int(seed) * 100003 + i))
    return out


def dipoles(shape, n_dipoles: int, seed: int) -> List[Tuple[Tuple[int, int, int], float]]:
    """Distinct interior cell pairs (+1, -1) per dipole."""
    D, H, W = shape
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_dipoles):
        while True:
            p = tuple(int(v) for v in (rng.integers(1, D - 1), rng.integers(1, H - 1), rng.integers(1, W - 1)))
            q = tuple(int(v) for v in (rng.integers(1, D - 1), rng.integers(1, H - 1), rng.integers(1, W - 1)))
            if p != q:
                break
        out += [(p, 1.0), (q, -1.0)]
    return out


def device_fields(shape, sigma_ln: float, n_dipoles: int, seed: int, device):
    """(sx, sy, sz, b) FP32 tensors on ``device``."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    sig = []
    for _ in range(3):
        z = torch.randn(shape, generator=g, device=device, dtype=torch.float32)
        sig.append(torch.exp(z * float(sigma_ln)))
    b = torch.zeros(shape, device=device, dtype=torch.float32)
    for (p, v) in dipoles(shape, n_dipoles, seed):
        b[p] += v
    return sig[0], sig[1], sig[2], b


def host_fields(shape, sigma_ln: float, n_dipoles: int, seed: int):
    """(sx, sy, sz, b) FP32 host arrays from ``numpy.random.default_rng(seed)``.

    Draw order: sx, sy, sz (each a (D, H, W) standard-normal field, times
    ``sigma_ln``, exponentiated, rounded to FP32), then per dipole two
    distinct interior cells (indices 1 .. n-2 per axis, drawn z, y, x) that
    get +1 and -1."""
    D, H, W = shape
    rng = np.random.default_rng(int(seed))
    sig = [np.exp(float(sigma_ln) * rng.standard_normal(shape)).astype(np.float32)
           for _ in range(3)]
    b = np.zeros(shape, np.float32)
    for _ in range(int(n_dipoles)):
        while True:
            p = (int(rng.integers(1, D - 1)), int(rng.integers(1, H - 1)), int(rng.integers(1, W - 1)))
            q = (int(rng.integers(1, D - 1)), int(rng.integers(1, H - 1)), int(rng.integers(1, W - 1)))
            if p != q:
                break
        b[p] += 1.0
        b[q] -= 1.0
    return sig[0], sig[1], sig[2], b


def host_fields_many(specs, workers: int = 16):
    """``host_fields`` for every (shape, sigma_ln, n_dipoles, seed) of ``specs``,
    generated on ``workers`` host threads (numpy's generators release the GIL
    while they fill large arrays); returns a list in ``specs`` order."""
    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        return list(ex.map(lambda sp: host_fields(*sp), specs))


def random_modes_ic(N: int, seed: int, kmax: int = 8, urms: float = 1.0):
    """Divergence-free start (u, v) of the 2D periodic flow, ((N, N) FP32, [j, i]).

    A stream function at the cell corners (i h, j h), h = 2 pi / N, sums the
    modes 0 < |k| <= kmax over (ky >= 0, kx in -kmax..kmax, skipping ky = 0,
    kx <= 0) with amplitude |k|^-3 and a phase drawn uniformly in [0, 2 pi)
    from ``numpy.random.default_rng(seed)`` in (ky, kx) row-major order; u is
    its forward difference along j and v minus its forward difference along
    i (divergence-free on the staggered grid), scaled to RMS speed ``urms``."""
    h = 2.0 * math.pi / N
    x = np.arange(N) * h
    X, Y = np.meshgrid(x, x)
    rng = np.random.default_rng(int(seed))
    psi = np.zeros((N, N))
    for ky in range(kmax + 1):
        for kx in range(-kmax, kmax + 1):
            if ky == 0 and kx <= 0:
                continue
            k = math.hypot(kx, ky)
            if k > kmax:
                continue
            psi += k ** -3 * np.cos(kx * X + ky * Y + rng.uniform(0.0, 2.0 * math.pi))
    u = (np.roll(psi, -1, 0) - psi) / h
    v = -(np.roll(psi, -1, 1) - psi) / h
    s = urms / math.sqrt(float(np.mean(u * u + v * v)))
    return (u * s).astype(np.float32), (v * s).astype(np.float32)


def urban_blocks(shape, seed: int, cover: float = 0.2, foot=(8, 24), height=(5, 30)):
    """Obstacle mask of axis-aligned blocks standing on the ground plane (k = 0).

    Blocks with footprints drawn from U{foot} x U{foot} cells and heights
    U{height} are added at seeded positions until at least ``cover`` of the
    ground plane is built on.  Returns a bool array of ``shape`` (D, H, W).
    """
    D, H, W = shape
    rng = np.random.default_rng(seed)
    mask = np.zeros(shape, dtype=bool)
    ground = np.zeros((H, W), dtype=bool)
    while ground.mean() < cover:
        fy, fx = (int(rng.integers(foot[0], foot[1] + 1)) for _ in range(2))
        hz = int(rng.integers(height[0], height[1] + 1))
        y0 = int(rng.integers(1, max(2, H - fy - 1)))
        x0 = int(rng.integers(1, max(2, W - fx - 1)))
        mask[:min(hz, D - 1), y0:y0 + fy, x0:x0 + fx] = True
        ground[y0:y0 + fy, x0:x0 + fx] = True
    return mask


def log_law_profile(D: int, ustar: float = 0.5, z0: float = 0.1, kappa: float = 0.41, dz=1.0):
    """Inflow u(z) = (u*/kappa) ln((z + z0)/z0) at cell-centre heights z = (k + 1/2) dz."""
    z = (np.arange(D) + 0.5) * dz
    return (ustar / kappa * np.log((z + z0) / z0)).astype(np.float32)
