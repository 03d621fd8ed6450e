"""CPU restatement of the 2D periodic Kolmogorov-flow time step.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP library's
2D flow kernels are compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path never calls it and has no CPU fallback.

The configuration (256^2 periodic on [0, 2 pi)^2, nu = 1/100, f_x = sin 4y,
divergence-free random-mode initial velocity, explicit 5-point diffusion and
upwind advection stencils, a 6-level multigrid pressure projection to 1e-6
relative residual, dt = 0.005) has no counterpart code in the reference: its
folder holds an empty README.  The pieces follow the reference's 2D layers:
the explicit two-stage update ``x + K(x + K(x)/2)`` of
ARTn/Advection_diffusion.ipynb:291-296 (ARTs/advection_diffusion_2D.py:131-141),
the 2D mean restriction / nearest prolongation / Jacobi multigrid of
ARTn/Parallel_multigrid_2D.ipynb#c5-#c20, and the urban-flow order
"momentum predictor, divergence source, multigrid pressure, gradient
correction" of UM/AI_HFM_v3.ipynb cell 25.  What the reference lacks (periodic
wrap, velocity-signed upwinding, an exact projection) is this build's choice:

Discretisation (staggered MAC grid, h = 2 pi / N, arrays indexed [j, i]):
  u[j, i] at (x, y) = (i h, (j + 1/2) h)      (x-faces)
  v[j, i] at       ((i + 1/2) h, j h)         (y-faces)
  phi[j, i] at     ((i + 1/2) h, (j + 1/2) h) (cell centres)
  F_u = nu/h^2 lap5(u) - (u du + vbar du') / h + sin(4 y)
        du  = u - u_w if u > 0 else u_e - u;  du' = u - u_s if vbar > 0 else u_n - u
        vbar = (v[j,i-1] + v[j,i] + v[j+1,i-1] + v[j+1,i]) / 4
  F_v = nu/h^2 lap5(v) - (ubar dv + v dv') / h
        ubar = (u[j-1,i] + u[j-1,i+1] + u[j,i] + u[j,i+1]) / 4
  half step  u_h = u + dt/2 F(u);  full step u* = u + dt F(u_h)
  projection b = h (u*_i - u*_{i+1} + v*_j - v*_{j+1}) - mean,  A phi = b with
             A = 4 phi - (phi_e + phi_w + phi_n + phi_s)  (A = -h^2 div grad),
             u = u* - (phi - phi_w)/h,  v = v* - (phi - phi_s)/h,
  so div u = (b - A phi) / (-h^2): the divergence left is the solver residual.

Solver: FP64 conjugate gradients (phi and r in FP64, warm-started from the
previous step's phi) preconditioned by one FP32 V(2,2) cycle: weighted Jacobi
(omega = 4/5), 2x2 mean restriction, nearest prolongation with coarse
correction factor alpha, Galerkin coarse operators (the same stencil times
2^-level) and the exact (pseudo-inverse, FP64) solve of the 8x8 coarsest
level.  FP32 storage of phi cannot reach 1e-6: a smooth pressure is ~10^3
times its right-hand side, so rounding phi to FP32 leaves a residual of
~1e-3 of b.

Pinning (tests/test_oracle_flow2d.py): the laminar Kolmogorov state is a
fixed point of the step; a Taylor-Green vortex decays at exp(-2 nu t) to the
scheme's first-order accuracy; the projection leaves div u at the solver
tolerance; A is the 5-point Laplacian (Fourier symbols); the pseudo-inverse
solves the coarsest level exactly; the V-cycle is symmetric positive definite
on mean-zero vectors.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

f32 = np.float32


# ---------------------------------------------------------------------------
# Initial condition and forcing
# ---------------------------------------------------------------------------


def random_modes_ic(N: int, seed: int, kmax: int = 8, urms: float = 1.0):
    """Divergence-free (u, v) from a stream function of random Fourier modes.

    psi at the cell corners (i h, j h) is a sum over wavevectors 0 < |k| <= kmax
    (ky > 0, or ky = 0 and kx > 0) of |k|^-3 cos(k . x + theta_k), theta_k from
    ``numpy.random.default_rng(seed)`` drawn in (ky, kx) row-major order, so the
    velocity modes have amplitude ~|k|^-2.  u = d psi/dy, v = -d psi/dx by
    corner differences (discretely divergence-free), scaled to RMS speed
    ``urms``; FP32.
    """
    h = 2 * np.pi / N
    x = np.arange(N) * h
    X, Y = np.meshgrid(x, x)          # [j, i]
    rng = np.random.default_rng(seed)
    psi = np.zeros((N, N))
    for ky in range(0, kmax + 1):
        for kx in range(-kmax, kmax + 1):
            if ky == 0 and kx <= 0:
                continue
            k = math.hypot(kx, ky)
            if k > kmax:
                continue
            psi += k ** -3 * np.cos(kx * X + ky * Y + rng.uniform(0.0, 2 * np.pi))
    u = (np.roll(psi, -1, 0) - psi) / h
    v = -(np.roll(psi, -1, 1) - psi) / h
    s = urms / math.sqrt(float(np.mean(u * u + v * v)))
    return (u * s).astype(f32), (v * s).astype(f32)


def forcing_u(N: int, amp: float = 1.0, k: int = 4):
    """f_x = amp sin(k y) at the u points, y = (j + 1/2) h; shape (N, 1) FP32."""
    h = 2 * np.pi / N
    y = (np.arange(N) + 0.5) * h
    return (amp * np.sin(k * y)).astype(f32)[:, None]


# ---------------------------------------------------------------------------
# Momentum
# ---------------------------------------------------------------------------


def _r(a, dj, di):
    """a[j + dj, i + di] with periodic wrap."""
    return np.roll(a, (-dj, -di), axis=(0, 1))


def momentum_rhs(u, v, nu_h2, inv_h, fu):
    """(F_u, F_v) in FP32 with the kernel's operation order."""
    u, v = u.astype(f32), v.astype(f32)
    nu_h2, inv_h, q = f32(nu_h2), f32(inv_h), f32(0.25)
    ue, uw, un, us = _r(u, 0, 1), _r(u, 0, -1), _r(u, 1, 0), _r(u, -1, 0)
    ve, vw, vn, vs = _r(v, 0, 1), _r(v, 0, -1), _r(v, 1, 0), _r(v, -1, 0)
    vbar = q * ((vw + v) + (_r(v, 1, -1) + vn))
    ubar = q * ((us + _r(u, -1, 1)) + (u + ue))
    dux = np.where(u > 0, u - uw, ue - u)
    duy = np.where(vbar > 0, u - us, un - u)
    dvx = np.where(ubar > 0, v - vw, ve - v)
    dvy = np.where(v > 0, v - vs, vn - v)
    lap_u = ((ue + uw) + (un + us)) - f32(4) * u
    lap_v = ((ve + vw) + (vn + vs)) - f32(4) * v
    Fu = nu_h2 * lap_u - inv_h * (u * dux + vbar * duy) + fu
    Fv = nu_h2 * lap_v - inv_h * (ubar * dvx + v * dvy)
    return Fu.astype(f32), Fv.astype(f32)


def divergence_rhs(u, v, h):
    """b = h (u - u_e + v - v_n) in FP32, then the FP64 mean removed (FP64)."""
    b = (f32(h) * ((u - _r(u, 0, 1)) + (v - _r(v, 1, 0)))).astype(f32)
    b64 = b.astype(np.float64)
    return b64 - b64.sum() / b64.size


def apply_periodic(x, s=1.0):
    """A x = s (4 x - (x_e + x_w + x_n + x_s)) in the dtype of x."""
    t = x.dtype.type
    return t(s) * (t(4) * x - ((_r(x, 0, 1) + _r(x, 0, -1)) + (_r(x, 1, 0) + _r(x, -1, 0))))


# ---------------------------------------------------------------------------
# Multigrid (periodic, constant coefficient) and the FP64 CG projection
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class PeriodicMG:
    sizes: List[int]          # N, N/2, ..., coarsest
    pinv: np.ndarray          # FP64 pseudo-inverse of the coarsest operator
    omega: float
    alpha: float
    nu1: int
    nu2: int


def coarsest_pinv(n: int, scale: float) -> np.ndarray:
    """Pseudo-inverse of the n x n periodic operator ``scale * A``:
    inv(M + J/n^2) - J/n^2 with J the all-ones matrix (A's null space is the
    constant vector, on which J/n^2 is the identity)."""
    m = n * n
    M = np.zeros((m, m))
    for q in range(m):
        e = np.zeros(m)
        e[q] = 1.0
        M[:, q] = apply_periodic(e.reshape(n, n), scale).ravel()
    J = np.full((m, m), 1.0 / m)
    return np.linalg.inv(M + J) - J


def build_mg(N: int, coarsest: int = 8, omega: float = 0.8, alpha: float = 1.8,
             nu1: int = 2, nu2: int = 2) -> PeriodicMG:
    if N & (N - 1) or N < coarsest * 2:
        raise ValueError("N must be a power of two above the coarsest size")
    sizes = [N]
    while sizes[-1] > coarsest:
        sizes.append(sizes[-1] // 2)
    return PeriodicMG(sizes, coarsest_pinv(sizes[-1], 2.0 ** -(len(sizes) - 1)),
                      omega, alpha, nu1, nu2)


def _jacobi(x, b, s, omega):
    """FP32 weighted Jacobi x + omega (b - s A x) / (4 s); x = None: zero guess."""
    d = f32(4 * s)
    if x is None:
        return (f32(omega) * b / d).astype(f32)
    return (x + f32(omega) * (b - apply_periodic(x, s)) / d).astype(f32)


def _restrict(r):
    n = r.shape[0] // 2
    return (f32(0.25) * ((r[0::2, 0::2] + r[0::2, 1::2]) + (r[1::2, 0::2] + r[1::2, 1::2]))).astype(f32)


def _prolong(e):
    return np.repeat(np.repeat(e, 2, 0), 2, 1)


def vcycle(mg: PeriodicMG, b, level: int = 0):
    """z = V(b) in FP32 from a zero guess; level l's operator is 2^-l A."""
    s = 2.0 ** -level
    if level == len(mg.sizes) - 1:
        return (mg.pinv @ b.astype(np.float64).ravel()).reshape(b.shape).astype(f32)
    x = _jacobi(None, b, s, mg.omega)
    for _ in range(mg.nu1 - 1):
        x = _jacobi(x, b, s, mg.omega)
    bc = _restrict((b - apply_periodic(x, s)).astype(f32))
    ec = vcycle(mg, bc, level + 1)
    x = (x + f32(mg.alpha) * _prolong(ec)).astype(f32)
    for _ in range(mg.nu2):
        x = _jacobi(x, b, s, mg.omega)
    return x


@dataclasses.dataclass
class ProjectionResult:
    phi: np.ndarray          # FP64
    iters: int
    rel: float
    status: str


def project_solve(mg: PeriodicMG, b64, phi0, tol: float = 1e-6, maxit: int = 100):
    """FP64 CG on A phi = b from ``phi0``, preconditioned by one FP32 V-cycle.

    r = b - A phi (FP64), rl = f32(r); per iteration z = V(rl), rz = rl . z,
    beta = rz / rz_prev (0 first), p = f32(beta p + z), q = A p (FP64),
    a = rz / p.q, phi += a p, r -= a q; stop when ||r|| <= tol ||b||.
    """
    phi = np.array(phi0, dtype=np.float64)
    r = b64 - apply_periodic(phi)
    bn = math.sqrt(float(np.sum(b64 * b64)))
    rr = float(np.sum(r * r))
    if bn == 0.0:
        return ProjectionResult(phi, 0, 0.0, "converged")
    if math.sqrt(rr) <= tol * bn:
        return ProjectionResult(phi, 0, math.sqrt(rr) / bn, "converged")
    rz, p = 0.0, None
    for it in range(maxit):
        rl = r.astype(f32)
        z = vcycle(mg, rl)
        rzn = float(np.sum(rl.astype(np.float64) * z.astype(np.float64)))
        beta = 0.0 if it == 0 else rzn / rz
        rz = rzn
        p = z if beta == 0.0 else (np.float64(f32(beta)) * p + z).astype(f32)
        q = apply_periodic(p.astype(np.float64))
        pq = float(np.sum(p.astype(np.float64) * q))
        if not (pq > 0.0) or not math.isfinite(pq):
            return ProjectionResult(phi, it, math.sqrt(rr) / bn, "breakdown")
        a = rz / pq
        phi += a * p.astype(np.float64)
        r -= a * q
        rr = float(np.sum(r * r))
        rel = math.sqrt(rr) / bn
        if rel <= tol:
            return ProjectionResult(phi, it + 1, rel, "converged")
    return ProjectionResult(phi, maxit, math.sqrt(rr) / bn, "maxit")


# ---------------------------------------------------------------------------
# Time step
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class FlowState:
    u: np.ndarray            # FP32 [N, N]
    v: np.ndarray
    phi: np.ndarray          # FP64 projection potential (dt * pressure)
    iters: List[int] = dataclasses.field(default_factory=list)


def substeps(u, v, dt: float, h: float, cfl_max: float = 0.5) -> int:
    """Substeps of one step: ceil((max|u| + max|v|) dt / h / cfl_max), at least 1,
    in FP32 with the kernel's order (coefficient dt / (h cfl_max) rounded once)."""
    coef = f32(dt / h / cfl_max)
    q = f32(f32(np.abs(u).max()) + f32(np.abs(v).max())) * coef
    return max(1, int(math.ceil(float(f32(q)))))


def substep(state: FlowState, mg: PeriodicMG, dt_s, nu: float, fu, tol: float = 1e-6):
    """Two-stage momentum over ``dt_s`` (FP32), divergence source, projection, correction."""
    N = state.u.shape[0]
    h = 2 * np.pi / N
    nu_h2, inv_h = f32(nu / h ** 2), f32(1.0 / h)
    dt_s = f32(dt_s)
    u, v = state.u, state.v
    Fu, Fv = momentum_rhs(u, v, nu_h2, inv_h, fu)
    uh = (u + f32(0.5) * dt_s * Fu).astype(f32)
    vh = (v + f32(0.5) * dt_s * Fv).astype(f32)
    Fu, Fv = momentum_rhs(uh, vh, nu_h2, inv_h, fu)
    us = (u + dt_s * Fu).astype(f32)
    vs = (v + dt_s * Fv).astype(f32)
    b = divergence_rhs(us, vs, h)
    res = project_solve(mg, b, state.phi, tol)
    phi = res.phi
    u_new = (us.astype(np.float64) - (phi - _r(phi, 0, -1)) * np.float64(inv_h)).astype(f32)
    v_new = (vs.astype(np.float64) - (phi - _r(phi, -1, 0)) * np.float64(inv_h)).astype(f32)
    state.u, state.v, state.phi = u_new, v_new, phi
    state.iters.append(res.iters)
    return res


def step(state: FlowState, mg: PeriodicMG, dt: float, nu: float, fu, tol: float = 1e-6,
         cfl_max: float = 0.5):
    """One step of ``dt``: CFL-limited substeps (``substeps``), each a full
    momentum + projection substep of dt / n.  Returns (n, last projection)."""
    N = state.u.shape[0]
    n = substeps(state.u, state.v, dt, 2 * np.pi / N, cfl_max)
    dt_s = f32(f32(dt) / f32(n))
    res = None
    for _ in range(n):
        res = substep(state, mg, dt_s, nu, fu, tol)
        if res.status != "converged":
            break
    return n, res


def divergence(u, v, h):
    """Cell divergence (u_e - u + v_n - v) / h in FP64."""
    u, v = u.astype(np.float64), v.astype(np.float64)
    return ((_r(u, 0, 1) - u) + (_r(v, 1, 0) - v)) / h


def laminar_state(N: int, nu: float, amp: float = 1.0, k: int = 4):
    """The discrete laminar Kolmogorov flow: u = amp sin(k y) / (nu lambda_h),
    lambda_h = (2 - 2 cos(k h)) / h^2 (the 5-point symbol), v = 0."""
    h = 2 * np.pi / N
    lam = (2 - 2 * math.cos(k * h)) / h ** 2
    y = (np.arange(N) + 0.5) * h
    u = np.repeat((amp * np.sin(k * y) / (nu * lam))[:, None], N, axis=1)
    return u.astype(f32), np.zeros((N, N), f32)


def taylor_green(N: int, amp: float = 1.0):
    """u = amp sin x cos y, v = -amp cos x sin y at the MAC points (FP32)."""
    h = 2 * np.pi / N
    i = np.arange(N)
    xf, xc = i * h, (i + 0.5) * h
    u = amp * np.sin(xf)[None, :] * np.cos(xc)[:, None]
    v = -amp * np.cos(xc)[None, :] * np.sin(xf)[:, None]
    return u.astype(f32), v.astype(f32)
