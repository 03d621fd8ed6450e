"""Time-step drivers on the libhfm.so kernels.

UrbanFlow follows the reference's urban-flow loop (UM/AI_HFM_v3.ipynb cell 25;
script twin ARTs/naiver_stokes_3D.py:571-727) step for step:

  velocity / pressure BCs
  -> momentum predictor  b = phi + 0.5 * (dif(phi) - u xadv(phi) - v yadv(phi) - w zadv(phi))
  -> velocity BCs on b
  -> momentum corrector  phi += dif(b_phi) - b_u xadv(b_phi) - b_v yadv(b_phi) - b_w zadv(b_phi)
  -> scalar transport + indicator BCs
  -> bluff body  f / (1 + dt sigma)
  -> u -= pu(p), v -= pv(p), w -= pw(p); velocity BCs
  -> b = -(Su(u) + Sv(v) + Sw(w)); source BC
  -> multi_itr sawtooth cycles of the 27-point Poisson layer (outlet zeroed;
     the prolonged fine correction w_512 is applied with no fine-level sweep)
  -> pressure BCs; gradient correction; velocity BCs; bluff body

Each arrow is one fused kernel launch (the three momentum components share
one 27-point advection pass; the divergence is one pass over u, v, w).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import layers as Ly

_P27 = np.array([[[2, 3, 2], [3, 6, 3], [2, 3, 2]],
                 [[3, 6, 3], [6, -88, 6], [3, 6, 3]],
                 [[2, 3, 2], [3, 6, 3], [2, 3, 2]]], dtype=np.float64) / 26.0
_PLANE = np.array([[0.014, 0.056, 0.014], [0.056, 0.22, 0.056], [0.014, 0.056, 0.014]])


def _dx_kernel():
    K = np.zeros((3, 3, 3)); K[:, :, 0] = -_PLANE; K[:, :, 2] = _PLANE
    return K


def _dy_kernel():
    K = np.zeros((3, 3, 3)); K[:, 0, :] = _PLANE; K[:, 2, :] = -_PLANE
    return K


def _dz_kernel():
    K = np.zeros((3, 3, 3)); K[0] = _PLANE; K[2] = -_PLANE
    return K


def urban_kernels(dt: float, dx: float, Re: float):
    """The reference's fixed weights (UM/AI_HFM_v3.ipynb cells 7-10): w1 diffusion,
    w2/w3/w4 advection and gradient (x dt/dx x 0.5), w5 Poisson, w6/w7/w8 divergence
    (x 0.5/(dx dt)).  The y/z kernels differentiate toward decreasing index."""
    s_adv = dt / dx * 0.5
    s_div = 0.5 / (dx * dt)
    return {
        "dif": _P27 * dt * Re / dx ** 2,
        "adv": (_dx_kernel() * s_adv, _dy_kernel() * s_adv, _dz_kernel() * s_adv),
        "div": (_dx_kernel() * s_div, _dy_kernel() * s_div, _dz_kernel() * s_div),
        "poisson": -_P27 / dx ** 2,
    }


class UrbanFlow:
    """The reference urban-flow time step on (1, D, H, W, 1) or (D, H, W) GPU fields."""

    def __init__(self, shape, sigma: Optional[torch.Tensor] = None, dt: float = 0.4,
                 dx: float = 1.0, Re: float = 1 / 20, ub: float = 1.0, multi_itr: int = 5,
                 lscalar: bool = True, device=None, coarsest: int = 1, j_itr: int = 1,
                 fused: bool = True):
        """``coarsest``: the pressure hierarchy halves every axis until the
        shortest is 1 (UM/AI_HFM_v3.ipynb:2571-2584) or 2 (the script twin's
        2^3 coarsest level, ARTs/naiver_stokes_3D.py:648-662); ``j_itr``:
        Jacobi sweeps per level below the finest (UM/AI_HFM_v3.ipynb:780).
        ``fused``: the pressure fills, obstacle dampings and the scalar's fill
        ride in the gradient and advection passes (bitwise the layer-by-layer
        step, which ``fused=False`` runs)."""
        self.shape = tuple(int(v) for v in shape)
        if len(self.shape) != 3 or min(self.shape) < 2:
            raise ValueError("UrbanFlow needs a (D, H, W) box with every edge >= 2")
        self.device = torch.device(device or "cuda")
        self.dt, self.ub, self.multi_itr, self.lscalar = float(dt), float(ub), int(multi_itr), lscalar
        k = urban_kernels(dt, dx, Re)
        self.k_mom = (k["dif"], *k["adv"])
        self.k_div = k["div"]
        self.k_grad = k["adv"]
        self.fused = bool(fused)
        self.grad = [Ly.FixedConv(K) for K in k["adv"]]
        if int(coarsest) not in (1, 2):
            raise ValueError("coarsest must be 1 or 2")
        shapes = [self.shape]
        while min(shapes[-1]) > int(coarsest):
            shapes.append(tuple(n // 2 for n in shapes[-1]))
        for s_f, s_c in zip(shapes, shapes[1:]):
            if any(a != 2 * b for a, b in zip(s_f, s_c)):
                raise ValueError(f"grid {self.shape} does not halve evenly down to its coarsest level")
        self.mg = Ly.SawtoothMG([Ly.FixedConv(k["poisson"], s) for s in shapes], zero_outlet=True,
                                fine_sweep=False, j_itr=j_itr)
        if sigma is None:
            sigma = torch.zeros(self.shape, device=self.device)
        self.sigma = sigma.reshape(self.shape).to(device=self.device, dtype=torch.float32).contiguous()
        self._last_wc = None    # the last cycle's correction one level below the finest

    # the scalar is damped with the velocities (the INHALE demo damps the velocities only)
    damp_scalar = True

    def _bluff(self, u, v, w, alpha):
        return Ly.bluff_body(u, v, w, alpha, self.sigma, dt=self.dt)

    def _bluff_all(self, u, v, w, alpha):
        if self.damp_scalar:
            return self._bluff(u, v, w, alpha)
        return (*self._bluff(u, v, w, None), alpha)

    def _inflow(self):
        return None

    def _bcv(self, u, v, w):
        return Ly.boundary_condition_velocity(u, v, w, ub=self.ub, inflow=self._inflow())

    def _scratch(self, u):
        D, H, W = self.shape
        n = 2 * D * W + 2 * H * W
        if getattr(self, "_bc_scratch", None) is None or self._bc_scratch.device != u.device:
            self._bc_scratch = torch.empty(n, dtype=torch.float32, device=u.device)
        return self._bc_scratch

    def _bcv_(self, u, v, w):
        """The velocity fill in place, on the step's own intermediates (it writes
        only the boundary planes; the step's inputs get the copying fill)."""
        return Ly.boundary_condition_velocity_(u, v, w, ub=self.ub, inflow=self._inflow(),
                                               scratch=self._scratch(u))

    def _bcv_bluff_(self, u, v, w):
        """The in-place fill and the damping, after ``gradient3(post_bluff=True)``."""
        return Ly.boundary_condition_velocity_bluff_(u, v, w, self.sigma, self.dt, ub=self.ub,
                                                     inflow=self._inflow(),
                                                     scratch=self._scratch(u))

    def _gradient(self, u, v, w, p):
        return (self.grad[0].subtract_from(u, p), self.grad[1].subtract_from(v, p),
                self.grad[2].subtract_from(w, p))

    def _momentum(self, u, v, w):
        """The predictor and corrector (UM/AI_HFM_v3.ipynb cell 25)."""
        bu, bv, bw = Ly.advect((u, v, w), (u, v, w), (u, v, w), 0.5, self.k_mom)
        bu, bv, bw = self._bcv_(bu, bv, bw)
        return Ly.advect((bu, bv, bw), (bu, bv, bw), (u, v, w), 1.0, self.k_mom)

    def _solve_pressure(self, p, b):
        return self._pressure_cycles(p, b)

    def step(self, u, v, w, p, alpha, donate: bool = False):
        """One time step; returns (u, v, w, p, alpha) as new tensors.  With
        ``donate`` the caller gives up u, v and w (as the reference's loop does,
        reassigning them): the input velocity fill then writes their boundary
        planes in place instead of copying the three fields, with bitwise the
        same results."""
        shape = tuple(u.shape)
        u, v, w, p, alpha = (t.reshape(self.shape).float().contiguous() for t in (u, v, w, p, alpha))
        distinct = len({t.data_ptr() for t in (u, v, w)}) == 3
        u, v, w = self._bcv_(u, v, w) if donate and distinct else self._bcv(u, v, w)
        out = self._step_fused(u, v, w, p, alpha) if self.fused else \
            self._step_layers(u, v, w, p, alpha)
        return tuple(t.reshape(shape) for t in out)

    def _step_layers(self, u, v, w, p, alpha):
        """The step layer by layer, in the reference's order (after the inputs'
        velocity fill)."""
        p = Ly.boundary_condition_pressure(p)
        u, v, w = self._momentum(u, v, w)
        if self.lscalar:
            (alpha,) = Ly.advect((u, v, w), (alpha,), (alpha,), 1.0, self.k_mom)
            alpha = Ly.boundary_condition_indicator(alpha)
        u, v, w, alpha = self._bluff_all(u, v, w, alpha)
        u, v, w = self._gradient(u, v, w, p)
        u, v, w = self._bcv_(u, v, w)
        b = Ly.conv_sum((u, v, w), self.k_div, scale=-1.0, source_bc=True)
        p = self._solve_pressure(p, b)
        p = Ly.boundary_condition_pressure(p)
        u, v, w = self._gradient(u, v, w, p)
        u, v, w = self._bcv_(u, v, w)
        u, v, w, alpha = self._bluff_all(u, v, w, alpha)
        return u, v, w, p, alpha

    def _step_fused(self, u, v, w, p, alpha):
        """The same step with the standalone passes folded into their neighbours:
        the scalar's advection also fills its faces and applies both of the
        step's dampings (alpha is not used between them); the first gradient
        reads p through the pressure fill (and writes the filled p) and damps
        the velocities before subtracting; the second fills p the same way and
        damps its result, which the in-place velocity fill completes."""
        u, v, w = self._momentum(u, v, w)
        nb = 2 if self.damp_scalar else 0
        if self.lscalar:
            (alpha,) = Ly.advect((u, v, w), (alpha,), (alpha,), 1.0, self.k_mom, indicator=True,
                                 sigma=self.sigma, dt=self.dt, nbluff=nb)
        elif nb:
            for _ in range(nb):   # the input scalar, damped with the velocities
                (alpha,) = Ly.bluff_body(alpha, None, None, None, self.sigma, dt=self.dt)
        u, v, w, p = Ly.gradient3(p, (u, v, w), self.k_grad, self.sigma, self.dt, p_bc=True,
                                  pre_bluff=True, out=(u, v, w))
        u, v, w = self._bcv_(u, v, w)
        b = Ly.conv_sum((u, v, w), self.k_div, scale=-1.0, source_bc=True)
        p = self._solve_pressure(p, b)
        u, v, w, p = Ly.gradient3(p, (u, v, w), self.k_grad, self.sigma, self.dt, p_bc=True,
                                  post_bluff=True, out=(u, v, w))
        u, v, w = self._bcv_bluff_(u, v, w)
        return u, v, w, p, alpha

    def run(self, u, v, w, p, alpha, ntime: int, writer=None, nrestart: int = 0):
        """The reference's time loop (UM/AI_HFM_v3.ipynb cell 25): steps
        itime = 1 .. ntime, each followed by ``writer.maybe_save(itime +
        nrestart, ...)`` (a ``snapshot.SnapshotWriter``; u, v, w, p and, with
        the scalar transport, alpha).  Returns the final fields."""
        for itime in range(1, int(ntime) + 1):
            # the caller's fields are left alone; later steps' inputs are the
            # loop's own
            u, v, w, p, alpha = self.step(u, v, w, p, alpha, donate=itime > 1)
            if writer is not None:
                extra = {"alpha": alpha} if self.lscalar else {}
                writer.maybe_save(itime + nrestart, u=u, v=v, w=w, p=p, **extra)
        return u, v, w, p, alpha

    def _pressure_cycles(self, p, b):
        rec = {}
        for _ in range(self.multi_itr):
            p = self.mg.cycle(p, b, record=rec)
        # the finest correction is the nearest prolongation of level 1's (or,
        # on a one-level hierarchy, level 0's own)
        self._last_wc = rec.get(1 if len(self.mg.ops) > 1 else 0)
        return p

    @property
    def last_w(self):
        """The last cycle's fine-level correction ``w_512 = CNN3D_prol_256(w_256)``
        (formed on request; the step applies it through the fused kernel)."""
        if self._last_wc is None or len(self.mg.ops) == 1:
            return self._last_wc
        return Ly.prolong_nearest(self._last_wc)

    def diverged(self, limit: float = 0.2) -> bool:
        """The reference's guard ``np.max(np.abs(w_512)) > 0.2`` (UM/AI_HFM_v3.ipynb:2625):
        w_512 repeats the coarse correction's values, so its maximum is read there."""
        return self._last_wc is not None and float(self._last_wc.abs().max()) > limit


class InhaleFlow(UrbanFlow):
    """The time step of ``UM/INHALE_demo.ipynb`` (cell 24): the urban step with
    the Petrov-Galerkin dissipation (cell 20) in place of the diffusion layer
    in both momentum stages, the bluff body on the velocities only, the scalar
    transport off by default (``LSCALAR = False``), dt = 0.5, ``eps_k`` =
    1e-3 (cell 22).  Same BCs and sawtooth pressure cycles as ``UrbanFlow``;
    the grid may be flat (the demo's nz = nx / 16: coarsest level 1 x r x r)."""

    def __init__(self, shape, sigma: Optional[torch.Tensor] = None, dt: float = 0.5,
                 dx: float = 1.0, Re: float = 1 / 20, ub: float = 1.0, multi_itr: int = 5,
                 eps_k: float = 1e-3, lscalar: bool = False, device=None, coarsest: int = 1,
                 j_itr: int = 1, fused: bool = True):
        super().__init__(shape, sigma=sigma, dt=dt, dx=dx, Re=Re, ub=ub, multi_itr=multi_itr,
                         lscalar=lscalar, device=device, coarsest=coarsest, j_itr=j_itr,
                         fused=fused)
        self.eps_k = float(eps_k)

    damp_scalar = False

    def _pg(self, u, v, w):
        return Ly.petrov_galerkin_dissipation(u, v, w, self.eps_k, self.sigma, self.dt,
                                              self.k_mom)

    def _momentum(self, u, v, w):
        """Both momentum stages with the Petrov-Galerkin dissipation (cell 24)."""
        adv = self.k_mom[1:]
        kx = self._pg(u, v, w)
        bu, bv, bw = Ly.advect_dissipative((u, v, w), (u, v, w), (u, v, w), kx, 0.5, adv)
        bu, bv, bw = self._bcv_(bu, bv, bw)
        kb = self._pg(bu, bv, bw)
        return Ly.advect_dissipative((bu, bv, bw), (bu, bv, bw), (u, v, w), kb, 1.0, adv)

    def diverged(self, limit: float = 0.5) -> bool:
        """The demo's guard: max |w_fine| of the last cycle above ``nsafe`` = 0.5
        (UM/INHALE_demo.ipynb:13478)."""
        return super().diverged(limit)


class AdvectionDiffusion2D:
    """The reference's 2D explicit advection-diffusion step
    (ARTn/Advection_diffusion.ipynb:291-296, ARTs/advection_diffusion_2D.py:131-141):
    one fixed 3x3 kernel K (diffusion and fixed-velocity advection folded in)
    and the two-stage update ``x <- x + K(x + K(x)/2)``, as two fused passes."""

    UPWIND = np.array([[0.0, 0.2, 0.0], [0.3, -1.0, 0.2], [0.0, 0.3, 0.0]])
    CENTRAL = np.array([[0.0, 0.15, 0.0], [0.25, -0.8, 0.15], [0.0, 0.25, 0.0]])

    def __init__(self, kernel=None):
        self.layer = Ly.FixedConv(self.UPWIND if kernel is None else kernel)

    def step(self, x):
        half = self.layer.axpy(x, x, 0.5)
        return self.layer.axpy(x, half, 1.0)


class UrbanProjection(UrbanFlow):
    """The urban-flow step with a 7-point multigrid pressure projection solved to a
    relative residual (the configuration's form of the step): log-law inflow on the
    x faces, free-slip top and no-slip ground (the reference's velocity BCs), the
    sigma-mask obstacle term, and instead of ``multi_itr`` sawtooth cycles of the
    27-point layer, ``A p = b`` with A the 7-point -Laplacian (Neumann faces, a
    Dirichlet outlet at x = W-1) solved by the FCG multigrid of ``hfm.mg`` in FP64
    with ``storage`` (FP16, BF16 or FP32) V-cycles, from zero, to ``tol``; ``outer``
    is the storage of the preconditioned residual and search direction.  The
    default is FP16: on this Neumann system BF16's 8-bit mantissa costs about
    2.5x the iterations of FP32, FP16's 11 bits about 1.1x, at BF16's bytes."""

    def __init__(self, shape, sigma=None, inflow=None, dt: float = 0.4, dx: float = 1.0,
                 Re: float = 1 / 20, storage: str = "f16", tol: float = 1e-8,
                 lscalar: bool = True, device=None, outer=None, tune: bool = True,
                 fused: bool = True):
        from .mg import MGBatch
        super().__init__(shape, sigma=sigma, dt=dt, dx=dx, Re=Re, ub=1.0, multi_itr=0,
                         lscalar=lscalar, device=device, fused=fused)
        D, H, W = self.shape
        self.inflow = None if inflow is None else torch.as_tensor(inflow, dtype=torch.float32).to(self.device)
        inv = 1.0 / dx ** 2
        one = torch.full(self.shape, inv, dtype=torch.float32)
        cx, cy, cz = one.clone(), one.clone(), one.clone()
        cx[:, :, 0] = 0.0       # Neumann at x = 0
        cy[:, 0, :] = 0.0       # Neumann at y = 0
        cz[0] = 0.0             # Neumann at the ground
        faces = (cx, cy, cz, torch.full((D, H), inv),   # Dirichlet outlet at x = W-1
                 torch.zeros((D, W)), torch.zeros((H, W)))
        self.solver = MGBatch([self.shape], storage=storage, tol=tol, maxit=500,
                              device=self.device, outer=outer)
        self.solver.set_faces(0, faces, torch.zeros(self.shape))
        self.solver.setup()
        self.tuning = None
        if tune:   # launch shapes and coarse-cycle threshold for this grid (~0.3 s)
            from .tune import tune as _tune
            self.tuning = _tune(self.solver)
        self.last_info = None

    def _inflow(self):
        return self.inflow

    def pressure(self, b):
        inst = self.solver.insts[0]
        inst.b32.copy_(b.reshape(-1))
        self.last_info = self.solver.solve()
        return inst.solution().float()

    def _solve_pressure(self, p, b):
        return self.pressure(b)
