"""PyTorch-CPU restatement of the config-mode solve, one array operation per step.

TEST INFRASTRUCTURE ONLY: the CPU baseline of ``bench.py`` (its
``cpu_baseline`` leg) and the tests that check it.  The product path never
imports it.

This is the unfused per-operation form the reference's Keras code runs: the
variable-coefficient operator as separate shifted products and sums
(``A_operator_<n>``: six face products, six shifts and the diagonal,
ART/BGS_multigrid_v1.ipynb#c40), the Jacobi sweep as separate array ops
(``w - A(w)/d + r/d``, #c42), restriction as ``avg_pool3d`` of the residual
and prolongation as ``repeat_interleave`` (ART/Anisotropic_3d_Multigrid_
7point.ipynb:194-284) -- each a separate full-array pass, in FP32 on the
host's cores (``torch.set_num_threads``).  The outer loop is the oracle's
FP64 flexible CG (``hfm_oracle.fcg_solve``), the hierarchy the oracle's
Galerkin operators (built in NumPy, outside the timed solve), the coarsest
level the oracle's dense inverse.

Grids whose levels halve evenly to the coarsest are supported (every edge of
the benchmark workload does); a ragged level raises ``ValueError``.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np
import torch
import torch.nn.functional as F

from . import hfm_oracle as O


class _Level:
    """One level's faces as FP32 tensors: the six per-cell face coefficients
    (low and high side of every axis) and the diagonal."""

    def __init__(self, op: O.FaceOp):
        ex, ey, ez = O._hi_faces(op)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32))  # noqa: E731
        self.cx, self.ex, self.cy, self.ey, self.cz, self.ez = (
            t(op.cx), t(ex), t(op.cy), t(ey), t(op.cz), t(ez))
        self.d = t(O.face_diag(op))
        self.shape = op.shape


def _shift(x: torch.Tensor, axis: int, step: int) -> torch.Tensor:
    """y[i] = x[i + step] along ``axis`` (step +1 or -1), zero outside the grid."""
    n = x.shape[axis]
    y = torch.zeros_like(x)
    if step > 0:
        y.narrow(axis, 0, n - 1).copy_(x.narrow(axis, 1, n - 1))
    else:
        y.narrow(axis, 1, n - 1).copy_(x.narrow(axis, 0, n - 1))
    return y


def apply(lv: _Level, x: torch.Tensor) -> torch.Tensor:
    """A x = d x - sum of face x neighbour products, as separate passes."""
    off = lv.cx * _shift(x, 2, -1)
    off = off + lv.ex * _shift(x, 2, 1)
    off = off + lv.cy * _shift(x, 1, -1)
    off = off + lv.ey * _shift(x, 1, 1)
    off = off + lv.cz * _shift(x, 0, -1)
    off = off + lv.ez * _shift(x, 0, 1)
    return lv.d * x - off


def _restrict(r: torch.Tensor) -> torch.Tensor:
    return F.avg_pool3d(r[None, None], 2)[0, 0]


def _prolong(e: torch.Tensor) -> torch.Tensor:
    return e.repeat_interleave(2, 0).repeat_interleave(2, 1).repeat_interleave(2, 2)


class Hierarchy:
    """The oracle's Galerkin hierarchy (FP32 operators) as torch levels."""

    def __init__(self, fine: O.FaceOp, max_coarse_edge: int = 4):
        h = O.build_hierarchy(fine, "f32", max_coarse_edge)
        shapes = [op.shape for op in h.ops]
        for a, b in zip(shapes, shapes[1:]):
            if any(p != 2 * q for p, q in zip(a, b)):
                raise ValueError(f"level {a} -> {b} does not halve evenly")
        self.levels: List[_Level] = [_Level(op) for op in h.ops]
        self.fine = _Level(h.fine32)
        self.coarse_inv = torch.from_numpy(h.coarse_inv.astype(np.float32))


def vcycle(h: Hierarchy, b: torch.Tensor, nu1=2, nu2=2, omega=2.0 / 3.0, alpha=1.6,
           level=0) -> torch.Tensor:
    lv = h.levels[level]
    if level == len(h.levels) - 1:
        return (h.coarse_inv @ b.reshape(-1)).reshape(b.shape)
    x = omega * b / lv.d
    for _ in range(nu1 - 1):
        x = x + omega * (b - apply(lv, x)) / lv.d
    bc = _restrict(b - apply(lv, x))
    ec = vcycle(h, bc, nu1, nu2, omega, alpha, level + 1)
    x = x + alpha * _prolong(ec)
    for _ in range(nu2):
        x = x + omega * (b - apply(lv, x)) / lv.d
    return x


def fcg_solve(sx, sy, sz, b, tol=1e-8, maxit=200, nu1=2, nu2=2, omega=2.0 / 3.0, alpha=1.6,
              threads: int = 0):
    """FP64 flexible CG preconditioned by one FP32 V(nu1, nu2) cycle, per op on
    the CPU.  Returns (x float64 ndarray, iterations, converged, solve seconds);
    the hierarchy is built before the clock starts."""
    import time
    if threads:
        torch.set_num_threads(int(threads))
    op = O.harmonic_faces(sx, sy, sz).astype("f32")
    h = Hierarchy(op)
    A = h.levels[0]
    b64 = torch.from_numpy(np.asarray(b, np.float32).astype(np.float64))
    t0 = time.perf_counter()
    x = torch.zeros_like(b64)
    r = b64.clone()
    bn = float(torch.linalg.vector_norm(b64))
    rz, p = 0.0, None
    it = 0
    ok = bn == 0.0
    while not ok and it < maxit:
        rl = r.float()
        z = vcycle(h, rl, nu1, nu2, omega, alpha).double()
        rzn = float((rl.double() * z).sum())
        p = z if it == 0 else z + (rzn / rz) * p
        rz = rzn
        q = _apply64(A, p)
        a = rz / float((p * q).sum())
        x = x + a * p
        r = r - a * q
        it += 1
        ok = float(torch.linalg.vector_norm(r)) / bn <= tol
    return x.numpy(), it, ok, time.perf_counter() - t0


def _apply64(lv: _Level, x: torch.Tensor) -> torch.Tensor:
    """A x in FP64 with the FP32 faces (the outer loop's exact operator)."""
    d = lambda a: a.double()  # noqa: E731
    off = d(lv.cx) * _shift(x, 2, -1)
    off = off + d(lv.ex) * _shift(x, 2, 1)
    off = off + d(lv.cy) * _shift(x, 1, -1)
    off = off + d(lv.ey) * _shift(x, 1, 1)
    off = off + d(lv.cz) * _shift(x, 0, -1)
    off = off + d(lv.ez) * _shift(x, 0, 1)
    diag = d(lv.cx) + d(lv.ex) + d(lv.cy) + d(lv.ey) + d(lv.cz) + d(lv.ez)
    return diag * x - off


def cell_iterations_per_s(specs, threads: int, make_fields) -> dict:
    """Solve each (shape, sigma_ln, n_dipoles, seed) of ``specs`` with ``threads``
    torch threads; returns the per-instance iterations, the solutions and the
    rate in fine-cell-iterations per second of solve time."""
    out, work, secs = [], 0, 0.0
    for shape, sig, nd, seed in specs:
        sx, sy, sz, b = make_fields(shape, sig, nd, seed)
        x, it, ok, t = fcg_solve(sx, sy, sz, b, threads=threads)
        out.append({"iters": it, "converged": ok, "x": x.reshape(shape), "t": t})
        work += math.prod(shape) * it
        secs += t
    return {"solves": out, "rate": work / secs if secs > 0 else 0.0, "seconds": secs}
