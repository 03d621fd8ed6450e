"""2D periodic Kolmogorov-flow driver on the libhfm.so flow kernel.

The configuration: an N x N periodic box [0, 2 pi)^2 (N = 256), nu = 1/100,
forcing f_x = sin(4 y), a divergence-free initial velocity from random
Fourier modes, explicit 5-point diffusion and upwind advection stencils with
the reference's two-stage update ``x + K(x + K(x)/2)``
(ARTn/Advection_diffusion.ipynb:291-296), and a pressure projection by a
6-level 2D multigrid (mean restriction, nearest prolongation, Jacobi;
ARTn/Parallel_multigrid_2D.ipynb#c5-#c20) to 1e-6 relative residual per
step, dt = 0.005.  The reference holds no code for this configuration (its
folder has an empty README); the step follows its urban-flow order
(UM/AI_HFM_v3.ipynb cell 25: momentum, divergence source, multigrid
pressure, gradient correction).

Discretisation and solver are those of ``hfm_flow2d_steps`` (include/hfm.h):
a staggered grid whose projection is exact (the divergence left is the
solver residual), CFL-limited substeps, FP64 CG over FP32 V-cycles.

``KolmogorovFlow`` advances ``nflows`` independent flows of one grid size;
each flow is one workgroup, so an ensemble fills the GPU.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
import torch

from . import _lib as L


def coarsest_pinv(n: int, scale: float) -> np.ndarray:
    """FP64 pseudo-inverse of ``scale * A`` on the n x n periodic level
    (A x = 4 x - sum of the 4 neighbours; null space: the constant vector),
    as inv(M + J) - J with J = ones / n^2.  Host set-up, computed once."""
    m = n * n
    M = np.zeros((m, m))
    idx = np.arange(m)
    j, i = idx // n, idx % n
    M[idx, idx] = 4.0 * scale
    for dj, di in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        nb = ((j + dj) % n) * n + (i + di) % n
        np.add.at(M, (idx, nb), -scale)
    J = np.full((m, m), 1.0 / m)
    return np.linalg.inv(M + J) - J


def forcing_profile(N: int, amp: float = 1.0, k: int = 4) -> np.ndarray:
    """f_x(y) = amp sin(k y) at the u points' heights y = (j + 1/2) h (FP32)."""
    h = 2 * math.pi / N
    return (amp * np.sin(k * (np.arange(N) + 0.5) * h)).astype(np.float32)


class KolmogorovFlow:
    """``nflows`` independent periodic flows on an N x N grid, advanced on the GPU."""

    def __init__(self, N: int = 256, nflows: int = 1, nu: float = 0.01, dt: float = 0.005,
                 amp: float = 1.0, k: int = 4, tol: float = 1e-6, nlev: int = 6,
                 omega: float = 0.8, alpha: float = 1.8, nu1: int = 2, nu2: int = 2,
                 maxit: int = 100, cfl_max: float = 0.5, max_sub: int = 64, team=None,
                 device=None):
        if N < 4 or N & (N - 1):
            raise ValueError("N must be a power of two >= 4")
        nc = N >> (nlev - 1)
        if nlev < 2 or nc < 2 or nc << (nlev - 1) != N or nc * nc > 1024:
            raise ValueError(f"{nlev} levels do not fit N = {N} (coarsest edge 2..32)")
        if nflows < 1:
            raise ValueError("nflows must be >= 1")
        # workgroups per flow: an ensemble fills the chip with one per flow; a
        # single flow spreads over a team (grids of 128^2 and more)
        if team is None:
            team = 16 if nflows == 1 and N >= 128 else 1
        if not 1 <= int(team) <= L.FLOW_TEAM_MAX:
            raise ValueError(f"team must be in 1..{L.FLOW_TEAM_MAX}")
        self.team = int(team)
        self.device = torch.device(device or "cuda")
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("KolmogorovFlow needs a GPU (the HIP library has no CPU path)")
        L.load()
        self.N, self.nflows, self.nlev = N, nflows, nlev
        self.nu, self.dt, self.tol = float(nu), float(dt), float(tol)
        h = 2 * math.pi / N
        self.h = h
        prm = np.zeros((), dtype=L.FLOW_PARAMS_DTYPE)
        prm["nu_h2"], prm["inv_h"], prm["h"], prm["dt"] = nu / h ** 2, 1.0 / h, h, dt
        prm["cfl_coef"] = dt / h / cfl_max
        prm["omega"], prm["alpha"] = omega, alpha
        prm["N"], prm["nu1"], prm["nu2"], prm["nlev"] = N, nu1, nu2, nlev
        prm["maxit"], prm["max_sub"], prm["tol"] = maxit, max_sub, tol
        prm["team"] = self.team
        self._prm = prm
        dev = self.device
        f32, f64 = torch.float32, torch.float64
        NN = N * N
        nmg = sum((N >> l) ** 2 for l in range(nlev))
        self.u = torch.zeros((nflows, N, N), dtype=f32, device=dev)
        self.v = torch.zeros_like(self.u)
        self.uh = torch.zeros_like(self.u)
        self.vh = torch.zeros_like(self.u)
        self.z = torch.zeros_like(self.u)
        self.p = torch.zeros_like(self.u)
        self.phi = torch.zeros((nflows, N, N), dtype=f64, device=dev)
        self.r = torch.zeros_like(self.phi)
        self.mgx = torch.zeros((nflows, nmg), dtype=f32, device=dev)
        self.mgy = torch.zeros_like(self.mgx)
        self.mgb = torch.zeros_like(self.mgx)
        self.fu = torch.from_numpy(np.tile(forcing_profile(N, amp, k), (nflows, 1))).to(dev)
        self.stats_buf = torch.zeros((nflows, L.FLOW_STATS.__len__()), dtype=f64, device=dev)
        self.team_buf = torch.zeros((nflows, L.FLOW_TEAM_BYTES // 8), dtype=f64, device=dev)
        self.pinv = torch.from_numpy(coarsest_pinv(nc, 2.0 ** -(nlev - 1))).to(dev)
        tab = np.zeros(nflows, dtype=L.FLOW_TASK_DTYPE)
        for q in range(nflows):
            for name in ("u", "v", "uh", "vh", "z", "p"):
                tab[name][q] = getattr(self, name)[q].data_ptr()
            for name in ("phi", "r"):
                tab[name][q] = getattr(self, name)[q].data_ptr()
            for name in ("mgx", "mgy", "mgb"):
                tab[name][q] = getattr(self, name)[q].data_ptr()
            tab["fu"][q] = self.fu[q].data_ptr()
            tab["stats"][q] = self.stats_buf[q].data_ptr()
            tab["team"][q] = self.team_buf[q].data_ptr()
        self.tasks = torch.from_numpy(tab.view(np.uint8).copy()).to(dev)
        assert self.u[0].numel() == NN

    # ------------------------------------------------------------------ state
    def set_state(self, q: int, u, v, phi=None):
        """Load flow ``q``'s velocity ((N, N) each, [j, i]) and projection potential."""
        self.u[q].copy_(torch.as_tensor(u, dtype=torch.float32))
        self.v[q].copy_(torch.as_tensor(v, dtype=torch.float32))
        if phi is None:
            self.phi[q].zero_()
        else:
            self.phi[q].copy_(torch.as_tensor(phi, dtype=torch.float64))
        self.stats_buf[q].zero_()

    def state(self, q: int):
        return self.u[q], self.v[q], self.phi[q]

    # ------------------------------------------------------------------ stepping
    def _launch(self, nsteps: int, flags: int):
        if self.team > 1:
            self.team_buf.zero_()      # barrier counters start at zero (a fill kernel)
        L.call("hfm_flow2d_steps", self.tasks.data_ptr(), self.nflows, self._prm.ctypes.data,
               self.pinv.data_ptr(), int(nsteps), flags, L.stream_handle(self.device))

    def step(self, nsteps: int = 1):
        """Advance every flow by ``nsteps`` steps of dt (one launch)."""
        self._launch(nsteps, 0)

    def fields(self) -> dict:
        """The state a restart needs: u, v and the projection potential as
        p ((N, N) each for one flow, (nflows, N, N) for an ensemble)."""
        if self.nflows == 1:
            return {"u": self.u[0], "v": self.v[0], "p": self.phi[0]}
        return {"u": self.u, "v": self.v, "p": self.phi}

    def run(self, nsteps: int, writer=None, nrestart: int = 0):
        """Advance ``nsteps`` steps; with a ``snapshot.SnapshotWriter`` the
        launches stop at every multiple of its ``n_out`` (counted from step
        ``nrestart``) and the fields are saved there.  Whole steps per
        launch, so the result equals ``step(nsteps)``."""
        if writer is None:
            self.step(nsteps)
            return
        done = 0
        while done < nsteps:
            k = min(nsteps - done, writer.n_out - (nrestart + done) % writer.n_out)
            self.step(k)
            done += k
            writer.maybe_save(nrestart + done, **self.fields())

    def restart(self, directory: str, nrestart: int):
        """Load the fields ``run`` saved at step ``nrestart`` (u, v, p files)."""
        from .snapshot import load_restart
        f, _ = load_restart(directory, nrestart, self.dt, names=("u", "v", "p"))
        u, v, p = (f[n].reshape(self.nflows, self.N, self.N) for n in ("u", "v", "p"))
        for q in range(self.nflows):
            self.set_state(q, u[q], v[q], p[q])

    def project(self):
        """Project the current velocities (no momentum stage)."""
        self._launch(0, L.FLOW_PROJECT_ONLY)

    def stats(self) -> np.ndarray:
        """Per-flow record (``_lib.FLOW_STATS`` fields) as a structured array."""
        a = self.stats_buf.cpu().numpy()
        out = np.zeros(self.nflows, dtype=[(n, "<f8") for n in L.FLOW_STATS])
        for q, n in enumerate(L.FLOW_STATS):
            out[n] = a[:, q]
        return out

    def divergence(self, q: int = 0) -> torch.Tensor:
        """Cell divergence (u_e - u + v_n - v) / h of flow ``q`` in FP64 (diagnostic)."""
        u, v = self.u[q].double(), self.v[q].double()
        return ((torch.roll(u, -1, 1) - u) + (torch.roll(v, -1, 0) - v)) / self.h

    def nbytes(self) -> int:
        ts = [self.u, self.v, self.uh, self.vh, self.z, self.p, self.phi, self.r, self.mgx,
              self.mgy, self.mgb, self.fu, self.stats_buf, self.team_buf]
        return sum(t.numel() * t.element_size() for t in ts)
