"""CPU restatement of the AI-HFM-Solver stencil + multigrid hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP library is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
calls it and has no CPU fallback.

Two halves:

* **Reference mode** -- the reference's Keras operators restated in NumPy with
  Keras semantics: ``Conv{1,2,3}D(kernel=3, stride=1, padding='SAME')`` is a
  cross-correlation with one zero ghost cell on every side, the 2^d "VALID"
  stride-2 restriction with weights ``1/2^d`` is a mean of children,
  ``UpSampling{1,2,3}D(2)`` is a nearest-neighbour repeat, and the smoother is
  ``w - A(w)/d + r/d``.  All functions are dimension-generic so the 1D and 2D
  golden vectors printed in the reference notebooks pin the same code that the
  3D path uses.

* **Config mode** -- the anisotropic resistivity forward solve
  ``div(sigma grad phi) = b`` with harmonic face coefficients, Galerkin
  (aggregation) coarse operators, V(nu1, nu2) weighted-Jacobi cycles with a
  coarse-correction over-relaxation factor, a dense FP64 coarsest solve and an
  FP64 defect-correction outer loop with emulated FP32 / BF16 / FP16 storage.
  The reference has no tolerance-driven solver; this half follows the solver
  configuration (64^3 log-normal anisotropic conductivity, dipole source,
  5 levels to 4^3, V(2,2), omega=2/3, mean restriction, nearest prolongation)
  and is pinned by the FP64 sparse direct solve and closed forms in the tests.

Pinning status (see the tests under tests/test_oracle_*.py):
  conv_same            golden (3D weight layout known answers), closed form,
                       library identity (torch conv3d)
  restrict/prolong     closed forms, R(P(x)) = x
  sawtooth_cycle       golden vectors (1D w_32/w_4, 2D w_2/w_4); the urban form
                       (no fine-level sweep) by a dense matrix restatement
  urban / INHALE steps dense cycle restatement, y-mirror equivariance, zero state
  2D explicit update   kernel sums, exact mass / drift / variance moments
  Multigrid_cnn cycle  dense matrix restatement (window-sum prolongation with
                       TensorFlow's after-the-end SAME padding); prolongation
                       against torch (interpolate + pad + conv2d)
  defect correction    the stationary iteration matrix (I - M A)^k
  var6 operator        reduces to the constant operator, dense brute force
  27-point layers      closed forms (linear fields, x^2)
  face operator        scipy.sparse assembly, sine eigenmodes, symmetry
  galerkin_coarsen     brute force P^T A P / 8
  vcycle / solve       FP64 sparse direct solve
  boundary functions   closed forms (plane constants and copies)
  Petrov-Galerkin      closed forms (constant / linear fields give no
  dissipation          dissipation; constant k reduces the flux to 3 k dif(f);
                       odd in the velocity; the cap 0.5/(1 + dt sigma))
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Sequence

import numpy as np

# ---------------------------------------------------------------------------
# Storage-precision emulation
# ---------------------------------------------------------------------------


def round_bf16(a) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even); returns float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def round_storage(a, storage: str) -> np.ndarray:
    """Round to the storage type ``storage`` in {"f64","f32","bf16","f16"}; result float64."""
    if storage == "f64":
        return np.asarray(a, dtype=np.float64)
    if storage == "f32":
        return np.asarray(a, dtype=np.float32).astype(np.float64)
    if storage == "bf16":
        return round_bf16(np.asarray(a, dtype=np.float32)).astype(np.float64)
    if storage == "f16":
        return np.asarray(a, dtype=np.float32).astype(np.float16).astype(np.float64)
    raise ValueError(f"unknown storage {storage!r}")


# ---------------------------------------------------------------------------
# Reference mode: Keras-semantics layers
# ---------------------------------------------------------------------------


def conv_same(x: np.ndarray, K: np.ndarray) -> np.ndarray:
    """Keras ConvNd(filters=1, kernel_size=3, strides=1, padding='SAME'), bias 0.

    Cross-correlation ``y[i] = sum_a K[a] x[i + a - 1]`` with zero ghosts.
    ``x`` has the spatial shape only (the reference's (1, ..., 1) batch and
    channel axes removed); ``K`` has shape (3,)*x.ndim and is the reference
    weight array ``w[0, ..., 0]`` (constant initializer fills the kernel in
    row-major order).  Follows ``ART/Anisotropic_3d_Multigrid_7point.ipynb:124-178``.
    """
    x = np.asarray(x)
    K = np.asarray(K)
    if K.shape != (3,) * x.ndim:
        raise ValueError(f"kernel shape {K.shape} does not match a {x.ndim}-d field")
    dt = np.result_type(x.dtype, np.float32)
    xp = np.pad(x.astype(dt), 1)
    y = np.zeros(x.shape, dtype=dt)
    for idx in np.ndindex(*K.shape):
        w = K[idx]
        if w == 0:
            continue
        sl = tuple(slice(a, a + n) for a, n in zip(idx, x.shape))
        y += dt.type(w) * xp[sl]
    return y


def restrict_mean(x: np.ndarray, factors: Optional[Sequence[int]] = None) -> np.ndarray:
    """Keras ConvNd(kernel=2, strides=2, 'VALID') with all weights 1/2^d: mean of children.

    Follows ``ART/Anisotropic_3d_Multigrid_7point.ipynb:194-235`` (w1 = 0.125) and
    the 2D/1D twins (0.25 / 0.5).  ``factors`` allows axes with factor 1.
    """
    x = np.asarray(x)
    f = tuple(factors) if factors is not None else (2,) * x.ndim
    shp = []
    for n, k in zip(x.shape, f):
        if n % k:
            raise ValueError("VALID restriction needs even extents")
        shp += [n // k, k]
    axes = tuple(range(1, 2 * x.ndim, 2))
    return x.reshape(shp).mean(axis=axes, dtype=np.float64).astype(x.dtype)


def prolong_nearest(x: np.ndarray, factors: Optional[Sequence[int]] = None) -> np.ndarray:
    """Keras UpSamplingNd(size=2): nearest-neighbour repeat (``…7point.ipynb:251-284``)."""
    x = np.asarray(x)
    f = tuple(factors) if factors is not None else (2,) * x.ndim
    for ax, k in enumerate(f):
        if k != 1:
            x = np.repeat(x, k, axis=ax)
    return x


def jacobi_ref(w, r, Aop: Callable, d):
    """One reference smoothing sweep ``w - A(w)/d + r/d`` (``…7point.ipynb:424-444``)."""
    return w - Aop(w) / d + r / d


def a7_const_kernel(Dx=200.0, Dy=100.0, Dz=50.0, dx=1.0, dy=1.0, dz=1.0) -> np.ndarray:
    """Constant-coefficient anisotropic 7-point weights w2[0,:,:,:,0].

    ``ART/Anisotropic_3d_Multigrid_7point.ipynb:75-90``: axis 0 carries Dz,
    axis 1 Dy, axis 2 Dx; centre ``2(Dx/dx^2+Dy/dy^2+Dz/dz^2)``; neighbours
    negative (the layer computes -Laplacian).
    """
    K = np.zeros((3, 3, 3))
    K[0, 1, 1] = K[2, 1, 1] = -Dz / dz ** 2
    K[1, 0, 1] = K[1, 2, 1] = -Dy / dy ** 2
    K[1, 1, 0] = K[1, 1, 2] = -Dx / dx ** 2
    K[1, 1, 1] = 2.0 * (Dx / dx ** 2 + Dy / dy ** 2 + Dz / dz ** 2)
    return K


def a7_source_kernel(dt=0.05, Dx=200.0, Dy=100.0, Dz=50.0, dx=1.0) -> np.ndarray:
    """Time-stepped variant of ``ARTn/Multigrid_3d_7point_source.ipynb:68-106`` (all ``*dt``)."""
    K = np.zeros((3, 3, 3))
    K[0, 1, 1] = K[2, 1, 1] = -dt * Dz / dx ** 2
    K[1, 0, 1] = K[1, 2, 1] = -dt * Dy / dx ** 2
    K[1, 1, 0] = K[1, 1, 2] = -dt * Dx / dx ** 2
    K[1, 1, 1] = 2.0 * dt * (Dx + Dy + Dz) / dx ** 2
    return K


# --- variable-coefficient operator of ART/BGS_multigrid_v1.ipynb (#c11, #c38, #c40)

VAR6_DIRS = ("xm1", "xp1", "ym1", "yp1", "zm1", "zp1")


def var6_kernels() -> dict:
    """The six 2-point difference kernels (BGS ``#c11``): ``x_i - x_nb``."""
    out = {}
    for name in VAR6_DIRS:
        K = np.zeros((3, 3, 3))
        K[1, 1, 1] = 1.0
        axis = {"x": 2, "y": 1, "z": 0}[name[0]]
        off = 0 if name[1] == "m" else 2
        idx = [1, 1, 1]
        idx[axis] = off
        K[tuple(idx)] = -1.0
        out[name] = K
    return out


def a_operator_var6(x: np.ndarray, mats: dict) -> np.ndarray:
    """``A_operator_n`` (BGS ``#c40``): sum over six directions of A_MAT_dir * (x - x_nb)."""
    ks = var6_kernels()
    y = np.zeros(x.shape, dtype=np.result_type(x.dtype, np.float32))
    for name in VAR6_DIRS:
        y = y + conv_same(x, ks[name]) * mats[name]
    return y


def var6_tot(mats: dict) -> np.ndarray:
    """``A_MAT_TOT`` (BGS ``#c38``): sum of the six coefficient fields."""
    return sum(mats[n] for n in VAR6_DIRS)


def restrict_var6(mats: dict) -> dict:
    """Coarse coefficient fields = mean restriction of every field (BGS ``#c38:16-33``)."""
    out = {n: restrict_mean(mats[n]) for n in VAR6_DIRS}
    out["tot"] = restrict_mean(mats["tot"]) if "tot" in mats else var6_tot(out)
    return out


def sawtooth_cycle(values, b, ops: List[Callable], diags: list, restrict=restrict_mean,
                   prolong=prolong_nearest, coarsest_sweeps: Optional[int] = None, bc=None,
                   record: Optional[dict] = None, zero_outlet: bool = False,
                   fine_sweep: bool = True, j_itr: int = 1):
    """One reference "sawtooth" multigrid cycle.

    ``…7point.ipynb:402-454`` / ``BGS#c42`` / ``ARTn/Parallel_multigrid_2D.ipynb#c20``:
    r = A(x) - b at the finest level, restrict r down the whole hierarchy,
    start from w = 0 at the coarsest level and do ``coarsest_sweeps`` sweeps
    there (default ``j_itr``; ``…_BC_Rec.ipynb:525`` does 2), then at each
    finer level prolong and do ``j_itr`` sweeps against that level's
    restricted residual, then ``x -= w``, one fine sweep on ``A x = b`` and the
    optional boundary fill.  ``ops[l]`` / ``diags[l]`` are the level-l
    operator and divisor, fine first.

    ``fine_sweep=False`` is the urban drivers' form: the finest level's
    correction is the prolonged coarse correction with no sweep of its own
    (``UM/AI_HFM_v3.ipynb:2610-2611``: ``w_512 = CNN3D_prol_256(w_256);
    values_p = values_p - w_512``; ``UM/INHALE_demo.ipynb:13454-13458``, where
    the fine sweep is commented out; ``ARTs/naiver_stokes_3D.py:696-699``).
    ``zero_outlet``: r[..., nx-1] = 0 after the residual and values[..., nx-1] = 0
    after the correction (``UM/AI_HFM_v3.ipynb:2574,2611-2612``).
    ``record[l]`` receives level l's correction (``record[0]`` the one that is
    subtracted from ``values``).
    """
    nlev = len(ops)
    if coarsest_sweeps is None:
        coarsest_sweeps = j_itr
    r = ops[0](values) - b
    if zero_outlet:
        r = zero_outlet_plane(r)
    rs = [r]
    for _ in range(1, nlev):
        rs.append(restrict(rs[-1]))
    w = np.zeros_like(rs[-1])
    for _ in range(coarsest_sweeps):
        w = jacobi_ref(w, rs[-1], ops[-1], diags[-1])
    if record is not None:
        record[nlev - 1] = w
    for lev in range(nlev - 2, -1, -1):
        wp = prolong(w)
        if record is not None:
            record[("prolonged", lev)] = wp
            record[("A_prolonged", lev)] = ops[lev](wp)
        w = wp
        if lev > 0 or fine_sweep:
            for _ in range(j_itr):
                w = jacobi_ref(w, rs[lev], ops[lev], diags[lev])
        if record is not None:
            record[lev] = w
    values = values - w
    if zero_outlet:
        values = zero_outlet_plane(values)
    values = values - ops[0](values) / diags[0] + b / diags[0]
    if bc is not None:
        values = bc(values)
    return values


def zero_outlet_plane(r):
    out = np.array(r, copy=True)
    out[..., -1] = 0
    return out


def boundary_condition_top(values: np.ndarray) -> np.ndarray:
    """Neumann copy of plane 1 into plane 0 along axis 0 (``…_BC.ipynb:340-344``)."""
    out = np.array(values, copy=True)
    out[0] = out[1]
    return out


# --- 27-point layers of UM/AI_HFM_v3.ipynb (cells 7-10) / ARTs/naiver_stokes_3D.py:37-153

_P27 = np.array([[[2, 3, 2], [3, 6, 3], [2, 3, 2]],
                 [[3, 6, 3], [6, -88, 6], [3, 6, 3]],
                 [[2, 3, 2], [3, 6, 3], [2, 3, 2]]], dtype=np.float64) / 26.0

_DX_PLANE = np.array([[0.014, 0.056, 0.014], [0.056, 0.22, 0.056], [0.014, 0.056, 0.014]])


def diffusion27_kernel(dt=0.4, Re=1 / 20, dx=1.0) -> np.ndarray:
    """``CNN3D_central_2nd_dif`` weights (``UM/AI_HFM_v3.ipynb`` cell 7)."""
    return _P27 * dt * Re / dx ** 2


def poisson27_kernel(dx=1.0) -> np.ndarray:
    """``CNN3D_A_*`` pressure weights w5 (``UM/AI_HFM_v3.ipynb`` cell 9)."""
    return -_P27 / dx ** 2


def central_x_kernel() -> np.ndarray:
    """Unscaled smoothed central x-derivative: - at x-1, + at x+1 (cell 8 ``p_div_x``)."""
    K = np.zeros((3, 3, 3))
    K[:, :, 0] = -_DX_PLANE
    K[:, :, 2] = _DX_PLANE
    return K


def central_y_kernel() -> np.ndarray:
    """Unscaled y-kernel: + at y-1, - at y+1 (cell 8 ``p_div_y``)."""
    K = np.zeros((3, 3, 3))
    K[:, 0, :] = _DX_PLANE
    K[:, 2, :] = -_DX_PLANE
    return K


def central_z_kernel() -> np.ndarray:
    """Unscaled z-kernel: + at z-1, - at z+1 (cell 8 ``p_div_z``)."""
    K = np.zeros((3, 3, 3))
    K[0] = _DX_PLANE
    K[2] = -_DX_PLANE
    return K


def advection_kernels(dt=0.4, dx=1.0, scale=0.5):
    """w2, w3, w4: advection/gradient layers (``× dt/dx × 0.5``, cell 8)."""
    s = dt / dx * scale
    return central_x_kernel() * s, central_y_kernel() * s, central_z_kernel() * s


def divergence_kernels(dt=0.4, dx=1.0, scale=0.5):
    """w6, w7, w8: divergence layers (``× 0.5/(dx dt)``, cell 10)."""
    s = scale / (dx * dt)
    return central_x_kernel() * s, central_y_kernel() * s, central_z_kernel() * s


def boundary_condition_velocity(u, v, w, ub=1.0):
    """``UM/AI_HFM_v3.ipynb`` cell 19 (generalised to a (D,H,W) box).

    Face assignments in the reference's order; every copy reads the *input*
    field (``tf.Variable(values_u)[...]``), later assignments overwrite earlier
    ones on shared edges.  Axis 2 = x: u=ub, v=w=0 at x=0 and x=nx-1.  Axis 1
    = y: u copies its neighbour plane, v=w=0.  Axis 0 = z: ground u=v=w=0; top
    u and v copy plane nz-2, w=0.
    """
    u0, v0, w0 = (np.asarray(a) for a in (u, v, w))
    u, v, w = (np.array(a, copy=True) for a in (u0, v0, w0))
    nz, ny, nx = u.shape
    u[:, :, 0] = ub; v[:, :, 0] = 0; w[:, :, 0] = 0
    u[:, :, nx - 1] = ub; v[:, :, nx - 1] = 0; w[:, :, nx - 1] = 0
    u[:, 0, :] = u0[:, 1, :]; v[:, 0, :] = 0; w[:, 0, :] = 0
    u[:, ny - 1, :] = u0[:, ny - 2, :]; v[:, ny - 1, :] = 0; w[:, ny - 1, :] = 0
    u[0] = 0; v[0] = 0; w[0] = 0
    u[nz - 1] = u0[nz - 2]; v[nz - 1] = v0[nz - 2]; w[nz - 1] = 0
    return u, v, w


def boundary_condition_pressure(p):
    """Cell 19: Dirichlet 0 at x=nx-1, Neumann copies (of the input) on the other faces."""
    p0 = np.asarray(p)
    p = np.array(p0, copy=True)
    nz, ny, nx = p.shape
    p[:, :, nx - 1] = 0
    p[:, :, 0] = p0[:, :, 1]
    p[:, 0, :] = p0[:, 1, :]
    p[:, ny - 1, :] = p0[:, ny - 2, :]
    p[0] = p0[1]
    p[nz - 1] = p0[nz - 2]
    return p


def boundary_condition_source(b):
    """Cell 19: the x=0 plane of the source copies the x=1 plane."""
    b0 = np.asarray(b)
    b = np.array(b0, copy=True)
    b[:, :, 0] = b0[:, :, 1]
    return b


def boundary_condition_indicator(a):
    """Cell 19: all six faces of the scalar set to 0."""
    a = np.array(a, copy=True)
    a[:, :, -1] = 0; a[:, :, 0] = 0; a[:, 0, :] = 0; a[:, -1, :] = 0; a[0] = 0; a[-1] = 0
    return a


def zero_outlet(r):
    """``r[:, :, nx-1] = 0`` after each MG residual / correction (``UM/AI_HFM_v3.ipynb:2574,2612``)."""
    r = np.array(r, copy=True)
    r[:, :, -1] = 0
    return r


def bluff_body(fields, sigma, dt):
    """Cell 19: implicit obstacle damping ``f / (1 + dt*sigma)`` for every field."""
    return tuple(f / (1.0 + dt * sigma) for f in fields)


# --- 2D explicit advection-diffusion (ARTn/Advection_diffusion.ipynb:73-94,291-296)

AD2D_UPWIND = np.array([[0.0, 0.2, 0.0], [0.3, -1.0, 0.2], [0.0, 0.3, 0.0]])
AD2D_CENTRAL = np.array([[0.0, 0.15, 0.0], [0.25, -0.8, 0.15], [0.0, 0.25, 0.0]])


def ad2d_step(x, K):
    """Two-stage update ``x <- x + K(x + K(x)/2)``."""
    return x + conv_same(x + 0.5 * conv_same(x, K), K)


def implicit_ad2d_kernel(dt=10.0, dx=1.0, dy=1.0, Dx=0.1, Dy=0.1, ub=1.0):
    """Implicit upwind AD stencil (``ARTn/Parallel_multigrid_2D.ipynb#c5``); sums to 1."""
    K = np.zeros((3, 3))
    K[0, 1] = -Dy * dt / dy ** 2
    K[1, 0] = -Dx * dt / dx ** 2 - ub * dt / dx
    K[1, 1] = 1 + 2 * (Dx * dt / dx ** 2 + Dy * dt / dy ** 2) + 2 * ub * dt / dx
    K[1, 2] = -Dx * dt / dx ** 2
    K[2, 1] = -Dy * dt / dy ** 2 - ub * dt / dx
    return K


def implicit_ad1d_kernel(dt=10.0, dx=1.0, Dx=0.1, ub=1.0):
    """1D twin (``ARTn/Parallel_multigrid_1D.ipynb#c4``); sums to 1."""
    return np.array([-ub * dt / dx - Dx * dt / dx ** 2,
                     1 + 2 * Dx * dt / dx ** 2 + ub * dt / dx,
                     -Dx * dt / dx ** 2])


# --- 2D multigrid of ARTn/Multigrid_cnn.ipynb (cells 2-21)

MGCNN_A = np.array([[0.0, -1.5, 0.0], [-51.5, 107.0, -1.5], [0.0, -51.5, 0.0]])
MGCNN_BIAS = float(np.float32(1.0 / 107.0))   # the bias layer's centre weight as stored


def prolong_upsum(x: np.ndarray, factors: Optional[Sequence[int]] = None) -> np.ndarray:
    """``UpSampling2D(2)`` then ``Conv2D(kernel 2x2 of ones, padding='SAME')``
    (``ARTn/Multigrid_cnn.ipynb:342-398``).  TensorFlow's SAME padding of an
    even window of 2 puts its one padding cell after the end, so output i sums
    upsampled cells i and i + 1 (0 past the end) on every upsampled axis."""
    f = tuple(factors) if factors is not None else (2,) * np.ndim(x)
    y = prolong_nearest(np.asarray(x, dtype=np.float64), f)
    for ax, k in enumerate(f):
        if k == 2:
            nxt = np.zeros_like(y)
            src = [slice(None)] * y.ndim
            dst = [slice(None)] * y.ndim
            src[ax] = slice(1, None)
            dst[ax] = slice(0, -1)
            nxt[tuple(dst)] = y[tuple(src)]
            y = y + nxt
    return y


def mgcnn_initial_condition(n: int = 128) -> np.ndarray:
    """Cells 14-16: a Gaussian (gamma = 10) times 0.8 on the interior plus a
    0.2 "blob" lattice (every 4th cell of every 4th row, and the cells off
    that lattice's columns in the other rows)."""
    x = -63.5 + 128.0 / n * np.arange(n)
    T = np.zeros((n, n))
    g2 = 2 * 10.0 ** 2
    for i in range(1, n - 1):
        for j in range(1, n - 1):
            T[i, j] = math.exp(-((x[i]) ** 2 + (x[j]) ** 2) / g2)
    blob = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if (i % 4 == 0 and j % 4 == 0) or (i % 4 != 0 and j % 4 != 0):
                blob[i, j] = 0.2
    return T * 0.8 + blob


def mgcnn_cycle(values, b, K=MGCNN_A, c=MGCNN_BIAS, j_itr: int = 2,
                record: Optional[dict] = None):
    """One cycle of ``ARTn/Multigrid_cnn.ipynb`` cell 21 (2D, levels down to 1 x 1).

    The smoother is ``w - bias(A(w)) + bias(r)`` with ``bias`` the layer that
    multiplies by its centre weight c = 1/107.  Order: a pre-sweep on A x = b;
    r = A x - b; r restricted (2x2 mean) to every level; from w = 0, ``j_itr``
    sweeps at the coarsest level; at each finer level but the finest, the
    upsampling prolongation (``prolong_upsum``) and ``j_itr`` sweeps; at the
    finest, the prolongation and one sweep against r; x -= w; a post-sweep
    on A x = b.  Returns (x, norm) with the notebook's residual monitor norm
    = sum(r) / (nx ny) of the cycle's r (``:605``)."""
    jac = lambda w, r: w - c * conv_same(w, K) + c * r  # noqa: E731
    values = np.asarray(values, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    values = jac(values, b)
    r = conv_same(values, K) - b
    rs = [r]
    while min(rs[-1].shape) > 1:
        rs.append(restrict_mean(rs[-1]))
    nlev = len(rs)
    w = np.zeros_like(rs[-1])
    for _ in range(j_itr):
        w = jac(w, rs[-1])
    if record is not None:
        record[nlev - 1] = w
    for lev in range(nlev - 2, 0, -1):
        w = prolong_upsum(w)
        for _ in range(j_itr):
            w = jac(w, rs[lev])
        if record is not None:
            record[lev] = w
    if nlev > 1:
        w = jac(prolong_upsum(w), r)
    if record is not None:
        record[0] = w
    values = values - w
    values = jac(values, b)
    return values, float(r.sum()) / r.size


def urban_levels(shape, coarsest: int = 1):
    """Reference-mode pressure hierarchy: halve every axis until the shortest
    is ``coarsest`` (1: ``UM/AI_HFM_v3.ipynb:2571-2584``, 512^3 down to 1^3;
    2: the script twin's ``w_2 = tf.zeros([1, 2, 2, 2, 1])`` coarsest level,
    ``ARTs/naiver_stokes_3D.py:648-662``)."""
    shapes = [tuple(shape)]
    while min(shapes[-1]) > coarsest:
        shapes.append(tuple(n // 2 for n in shapes[-1]))
    return shapes


def urban_pressure_cycles(p, b, multi_itr=5, dx=1.0, coarsest: int = 1, j_itr: int = 1,
                          record: Optional[dict] = None):
    """The urban drivers' ``multi_itr`` pressure cycles (``UM/AI_HFM_v3.ipynb:2570-2613``):
    27-point Poisson layer w5 at every level with its centre as divisor, outlet
    zeroing, ``j_itr`` sweeps per level below the finest and the prolonged
    correction applied with no fine-level sweep (``fine_sweep=False``)."""
    f32 = np.float32
    w5 = poisson27_kernel(dx)
    shapes = urban_levels(p.shape, coarsest)
    C = lambda x: conv_same(x.astype(f32), w5)  # noqa: E731
    ops = [C] * len(shapes)
    d = f32(w5[1, 1, 1])
    for _ in range(multi_itr):
        p = sawtooth_cycle(p, b, ops, [d] * len(shapes), zero_outlet=True, record=record,
                           fine_sweep=False, j_itr=j_itr)
    return p


def urban_step(u, v, w, p, alpha, sigma, dt=0.4, dx=1.0, Re=1 / 20, ub=1.0, multi_itr=5,
               lscalar=True, coarsest: int = 1, j_itr: int = 1):
    """One time step of ``UM/AI_HFM_v3.ipynb`` cell 25 (fields (D,H,W), float32 in/out).

    BCs; explicit two-stage momentum with the 27-point diffusion and smoothed
    central advection layers; scalar transport; bluff body; pressure-gradient
    correction; divergence source; ``multi_itr`` sawtooth cycles of the
    27-point Poisson layer with outlet zeroing and no fine-level sweep of the
    correction (``urban_pressure_cycles``); BCs, correction and bluff body
    again.  Returns (u, v, w, p, alpha, w_fine) with w_fine the last cycle's
    fine-level correction, the prolonged ``w_512`` (the reference's divergence
    guard reads max|w_512|, ``UM/AI_HFM_v3.ipynb:2625``).  ``coarsest`` /
    ``j_itr``: see ``urban_levels`` / ``sawtooth_cycle``.
    """
    f32 = np.float32
    dif = diffusion27_kernel(dt, Re, dx)
    ax, ay, az = advection_kernels(dt, dx)
    sx, sy, sz = divergence_kernels(dt, dx)
    C = lambda x, K: conv_same(x.astype(f32), K)  # noqa: E731

    def adv(vel, phi):
        uu, vv, ww = vel
        return ((C(phi, dif) - uu * C(phi, ax)) - vv * C(phi, ay)) - ww * C(phi, az)

    u, v, w = boundary_condition_velocity(u, v, w, ub=f32(ub))
    p = boundary_condition_pressure(p)
    a = [adv((u, v, w), phi) for phi in (u, v, w)]
    bu, bv, bw = (f32(0.5) * ai + phi for ai, phi in zip(a, (u, v, w)))
    bu, bv, bw = boundary_condition_velocity(bu, bv, bw, ub=f32(ub))
    c = [adv((bu, bv, bw), phi) for phi in (bu, bv, bw)]
    u, v, w = (ci + phi for ci, phi in zip(c, (u, v, w)))
    if lscalar:
        alpha = adv((u, v, w), alpha) + alpha
        alpha = boundary_condition_indicator(alpha)
    damp = f32(1.0) + f32(dt) * sigma
    u, v, w, alpha = (x / damp for x in (u, v, w, alpha))
    u, v, w = u - C(p, ax), v - C(p, ay), w - C(p, az)
    u, v, w = boundary_condition_velocity(u, v, w, ub=f32(ub))
    b = f32(-1.0) * ((C(u, sx) + C(v, sy)) + C(w, sz))
    b = boundary_condition_source(b)
    rec = {}
    p = urban_pressure_cycles(p, b, multi_itr, dx, coarsest, j_itr, record=rec)
    p = boundary_condition_pressure(p)
    u, v, w = u - C(p, ax), v - C(p, ay), w - C(p, az)
    u, v, w = boundary_condition_velocity(u, v, w, ub=f32(ub))
    u, v, w, alpha = (x / damp for x in (u, v, w, alpha))
    return u, v, w, p, alpha, rec.get(0)


def pg_coefficients(u, v, w, sigma, dt, eps_k, dif, ax, ay, az):
    """The diffusivities k_u, k_v, k_w of ``Petrov_Galerkin_dissipation``
    (``UM/INHALE_demo.ipynb`` cell 20, file lines 726-785), FP32:

      f~ = f / (1 + dt sigma)  for f in u, v, w
      k_f = 8*0.25*|1/3 (|u~| + |v~| + |w~|) dif(f~)|
            / (eps_k + (|xadv(f~)| + |yadv(f~)| + |zadv(f~)|)/3)
      k_f = min(k_f, 0.5) / (1 + dt sigma)

    Returns ((u~, v~, w~), (k_u, k_v, k_w))."""
    f32 = np.float32
    C = lambda x, K: conv_same(x.astype(f32), K)  # noqa: E731
    damp = f32(1.0) + f32(dt) * sigma
    ut, vt, wt = (x / damp for x in (u, v, w))
    speed = f32(1.0 / 3.0) * ((np.abs(ut) + np.abs(vt)) + np.abs(wt))
    ks = []
    for f in (ut, vt, wt):
        num = f32(2.0) * np.abs(speed * C(f, dif))
        den = f32(eps_k) + ((np.abs(C(f, ax)) + np.abs(C(f, ay))) + np.abs(C(f, az))) / f32(3.0)
        ks.append(np.minimum(num / den, f32(0.5)) / damp)
    return (ut, vt, wt), tuple(ks)


def pg_flux(f, k, dif):
    """``3*0.5*(k dif(f) + dif(f k) - f dif(k))`` of ``Petrov_Galerkin_dissipation``."""
    f32 = np.float32
    C = lambda x, K: conv_same(x.astype(f32), K)  # noqa: E731
    return f32(1.5) * ((k * C(f, dif) + C(f * k, dif)) - f * C(k, dif))


def pg_dissipation(u, v, w, sigma, dt, eps_k, dif, ax, ay, az):
    """(k_x, k_y, k_z) of ``Petrov_Galerkin_dissipation`` (INHALE demo cell 20)."""
    ft, ks = pg_coefficients(u, v, w, sigma, dt, eps_k, dif, ax, ay, az)
    return tuple(pg_flux(f, k, dif) for f, k in zip(ft, ks))


def inhale_step(u, v, w, p, alpha, sigma, dt=0.5, dx=1.0, Re=1 / 20, ub=1.0, multi_itr=5,
                eps_k=1e-3, lscalar=False, coarsest: int = 1, j_itr: int = 1):
    """One time step of ``UM/INHALE_demo.ipynb`` cell 24: the urban step with
    the Petrov-Galerkin dissipation in place of the diffusion layer in both
    momentum stages, the bluff body on the velocities only.  Returns
    (u, v, w, p, alpha, w_fine) as ``urban_step``."""
    f32 = np.float32
    dif = diffusion27_kernel(dt, Re, dx)
    ax, ay, az = advection_kernels(dt, dx)
    sx, sy, sz = divergence_kernels(dt, dx)
    C = lambda x, K: conv_same(x.astype(f32), K)  # noqa: E731

    def adv(vel, phi, first):
        uu, vv, ww = vel
        return ((first - uu * C(phi, ax)) - vv * C(phi, ay)) - ww * C(phi, az)

    u, v, w = boundary_condition_velocity(u, v, w, ub=f32(ub))
    p = boundary_condition_pressure(p)
    kx = pg_dissipation(u, v, w, sigma, dt, eps_k, dif, ax, ay, az)
    a = [adv((u, v, w), phi, k) for phi, k in zip((u, v, w), kx)]
    bu, bv, bw = (f32(0.5) * ai + phi for ai, phi in zip(a, (u, v, w)))
    bu, bv, bw = boundary_condition_velocity(bu, bv, bw, ub=f32(ub))
    kb = pg_dissipation(bu, bv, bw, sigma, dt, eps_k, dif, ax, ay, az)
    c = [adv((bu, bv, bw), phi, k) for phi, k in zip((bu, bv, bw), kb)]
    u, v, w = (ci + phi for ci, phi in zip(c, (u, v, w)))
    if lscalar:
        alpha = adv((u, v, w), alpha, C(alpha, dif)) + alpha
        alpha = boundary_condition_indicator(alpha)
    damp = f32(1.0) + f32(dt) * sigma
    u, v, w = (x / damp for x in (u, v, w))
    u, v, w = u - C(p, ax), v - C(p, ay), w - C(p, az)
    u, v, w = boundary_condition_velocity(u, v, w, ub=f32(ub))
    b = f32(-1.0) * ((C(u, sx) + C(v, sy)) + C(w, sz))
    b = boundary_condition_source(b)
    rec = {}
    p = urban_pressure_cycles(p, b, multi_itr, dx, coarsest, j_itr, record=rec)
    p = boundary_condition_pressure(p)
    u, v, w = u - C(p, ax), v - C(p, ay), w - C(p, az)
    u, v, w = boundary_condition_velocity(u, v, w, ub=f32(ub))
    u, v, w = (x / damp for x in (u, v, w))
    return u, v, w, p, alpha, rec.get(0)


# ---------------------------------------------------------------------------
# Config mode: face-coefficient operator, Galerkin hierarchy, V-cycles, solves
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class FaceOp:
    """7-point operator ``(A x)_c = sum_faces c_f (x_c - x_nb)`` with zero ghosts.

    ``cx[k,j,i]`` is the coefficient of the face between cells i-1 and i
    (i = 0: the low boundary face); ``bxh[k,j]`` is the high boundary face of
    cell i = W-1.  Likewise ``cy``/``byh`` (shape (D,W)) and ``cz``/``bzh``
    (shape (H,W)).  The diagonal is the sum of a cell's six faces.
    """
    cx: np.ndarray
    cy: np.ndarray
    cz: np.ndarray
    bxh: np.ndarray
    byh: np.ndarray
    bzh: np.ndarray

    @property
    def shape(self):
        return self.cx.shape

    def astype(self, storage: str) -> "FaceOp":
        return FaceOp(*(round_storage(a, storage) for a in dataclasses.astuple(self)))


def harmonic_faces(sx, sy, sz, h=1.0) -> FaceOp:
    """Face coefficients from per-cell conductivities (harmonic mean, / h^2).

    The boundary face of a cell gets the cell's own conductivity (the harmonic
    mean of a value with itself), with the zero ghost one full cell outside:
    the reference's SAME-padding Dirichlet convention.
    """
    sx, sy, sz = (np.asarray(s, dtype=np.float64) for s in (sx, sy, sz))
    inv = 1.0 / h ** 2

    def hm(a, b):
        return 2.0 * a * b / (a + b)

    cx = np.empty_like(sx); cx[:, :, 0] = sx[:, :, 0]; cx[:, :, 1:] = hm(sx[:, :, 1:], sx[:, :, :-1])
    cy = np.empty_like(sy); cy[:, 0, :] = sy[:, 0, :]; cy[:, 1:, :] = hm(sy[:, 1:, :], sy[:, :-1, :])
    cz = np.empty_like(sz); cz[0] = sz[0]; cz[1:] = hm(sz[1:], sz[:-1])
    return FaceOp(cx * inv, cy * inv, cz * inv,
                  sx[:, :, -1] * inv, sy[:, -1, :] * inv, sz[-1] * inv)


def _hi_faces(op: FaceOp):
    """Coefficient of every cell's +x, +y, +z face as full (D,H,W) arrays."""
    ex = np.concatenate([op.cx[:, :, 1:], op.bxh[:, :, None]], axis=2)
    ey = np.concatenate([op.cy[:, 1:, :], op.byh[:, None, :]], axis=1)
    ez = np.concatenate([op.cz[1:], op.bzh[None]], axis=0)
    return ex, ey, ez


def face_diag(op: FaceOp) -> np.ndarray:
    ex, ey, ez = _hi_faces(op)
    return op.cx + ex + op.cy + ey + op.cz + ez


def face_offdiag_sum(op: FaceOp, x: np.ndarray) -> np.ndarray:
    """``sum_nb c_f x_nb`` with zero ghosts."""
    ex, ey, ez = _hi_faces(op)
    s = np.zeros(x.shape, dtype=np.float64)
    s[:, :, 1:] += op.cx[:, :, 1:] * x[:, :, :-1]
    s[:, :, :-1] += ex[:, :, :-1] * x[:, :, 1:]
    s[:, 1:, :] += op.cy[:, 1:, :] * x[:, :-1, :]
    s[:, :-1, :] += ey[:, :-1, :] * x[:, 1:, :]
    s[1:] += op.cz[1:] * x[:-1]
    s[:-1] += ez[:-1] * x[1:]
    return s


def face_apply(op: FaceOp, x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    return face_diag(op) * x - face_offdiag_sum(op, x)


def face_sparse(op: FaceOp):
    """Assemble the operator as a scipy CSR matrix (cell index = (k*H + j)*W + i)."""
    import scipy.sparse as sp
    D, H, W = op.shape
    n = D * H * W
    idx = np.arange(n).reshape(D, H, W)
    ex, ey, ez = _hi_faces(op)
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [face_diag(op).ravel()]
    for c, a, b in ((op.cx[:, :, 1:], idx[:, :, 1:], idx[:, :, :-1]),
                    (ex[:, :, :-1], idx[:, :, :-1], idx[:, :, 1:]),
                    (op.cy[:, 1:, :], idx[:, 1:, :], idx[:, :-1, :]),
                    (ey[:, :-1, :], idx[:, :-1, :], idx[:, 1:, :]),
                    (op.cz[1:], idx[1:], idx[:-1]),
                    (ez[:-1], idx[:-1], idx[1:])):
        rows.append(a.ravel()); cols.append(b.ravel()); vals.append(-c.ravel())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def coarse_shape(shape):
    return tuple((n + 1) // 2 for n in shape)


def restrict_sum8(r: np.ndarray) -> np.ndarray:
    """R = P^T / 8: sum of (up to 2x2x2) children over 8; the mean for full aggregates."""
    D, H, W = r.shape
    Dc, Hc, Wc = coarse_shape(r.shape)
    rp = np.zeros((2 * Dc, 2 * Hc, 2 * Wc))
    rp[:D, :H, :W] = r
    return rp.reshape(Dc, 2, Hc, 2, Wc, 2).sum(axis=(1, 3, 5)) / 8.0


def prolong_agg(e: np.ndarray, fine_shape) -> np.ndarray:
    """Nearest-neighbour (piecewise-constant) prolongation onto ``fine_shape``."""
    D, H, W = fine_shape
    return np.repeat(np.repeat(np.repeat(e, 2, 0), 2, 1), 2, 2)[:D, :H, :W]


def galerkin_coarsen(op: FaceOp) -> FaceOp:
    """Coarse operator ``P^T A P / 8`` for 2x2x2 aggregates (ragged at odd edges).

    A coarse face coefficient is the sum of the fine face coefficients that
    cross it, over 8 (half the mean of the four fine faces for a full
    aggregate); faces interior to an aggregate cancel.
    """
    D, H, W = op.shape

    def agg2(a, axes):
        # sum over pairs along the listed axes (ragged tail kept as single)
        for ax in axes:
            n = a.shape[ax]
            nc = (n + 1) // 2
            pad = [(0, 0)] * a.ndim
            pad[ax] = (0, 2 * nc - n)
            a = np.pad(a, pad)
            shp = list(a.shape)
            shp[ax:ax + 1] = [nc, 2]
            a = a.reshape(shp).sum(axis=ax + 1)
        return a

    # faces at the low side of each aggregate along x: fine faces at even i
    cx = agg2(op.cx[:, :, 0::2], (0, 1)) / 8.0
    cy = agg2(op.cy[:, 0::2, :], (0, 2)) / 8.0
    cz = agg2(op.cz[0::2], (1, 2)) / 8.0
    bxh = agg2(op.bxh, (0, 1)) / 8.0
    byh = agg2(op.byh, (0, 1)) / 8.0
    bzh = agg2(op.bzh, (0, 1)) / 8.0
    return FaceOp(cx, cy, cz, bxh, byh, bzh)


def level_shapes(shape, max_coarse_edge: int = 4):
    """Hierarchy shapes: halve (rounding up) every axis while any edge exceeds ``max_coarse_edge``."""
    shapes = [tuple(shape)]
    while max(shapes[-1]) > max_coarse_edge:
        shapes.append(coarse_shape(shapes[-1]))
    return shapes


@dataclasses.dataclass
class Hierarchy:
    ops: List[FaceOp]          # per level, storage-rounded
    fine32: FaceOp             # fine operator as stored for the outer residual (FP32)
    coarse_inv: np.ndarray     # dense FP64 inverse of the coarsest operator
    storage: str


def build_hierarchy(fine: FaceOp, storage: str = "f32", max_coarse_edge: int = 4) -> Hierarchy:
    """Galerkin hierarchy.  Coarse operators are formed in FP64 from the FP32
    fine coefficients and then rounded to the storage type; the coarsest
    operator is inverted densely in FP64 from its FP32 coefficients."""
    f32 = fine.astype("f32")
    shapes = level_shapes(f32.shape, max_coarse_edge)
    exact = [f32]
    for _ in shapes[1:]:
        exact.append(galerkin_coarsen(exact[-1]).astype("f32"))
    ops = [o.astype(storage) for o in exact]
    Ac = face_sparse(exact[-1]).toarray()
    return Hierarchy(ops=ops, fine32=f32, coarse_inv=np.linalg.inv(Ac), storage=storage)


def jacobi_face(op: FaceOp, x, b, omega, storage):
    """Weighted Jacobi ``x + omega (b - A x)/d`` (x=None: zero initial guess)."""
    d = face_diag(op)
    if x is None:
        return round_storage(omega * b / d, storage)
    return round_storage(x + omega * (b - face_apply(op, x)) / d, storage)


def vcycle(h: Hierarchy, b, nu1=2, nu2=2, omega=2.0 / 3.0, alpha=1.0, level=0, x0=None,
           out_storage: Optional[str] = None):
    """One V(nu1, nu2) cycle on level ``level`` (``x0=None``: zero initial guess).

    ``out_storage`` rounds the cycle's result (its last post-sweep) to another
    type than the hierarchy's storage (the device's wider z buffer).

    Kernel order per non-coarsest level (each output rounded to storage):
      pre-sweep 1 (from zero when x0 is None), pre-sweeps 2..nu1, residual +
      restriction (b_c = R (b - A x)), recursive cycle on level+1 from zero,
      prolong + correct (x += alpha P e_c) fused into post-sweep 1,
      post-sweeps 2..nu2.  The coarsest level is solved exactly in FP64 and
      rounded to storage.
    """
    s = h.storage
    op = h.ops[level]
    if level == len(h.ops) - 1:
        e = (h.coarse_inv @ np.asarray(b, np.float64).ravel()).reshape(op.shape)
        return round_storage(e, s)
    x = jacobi_face(op, x0, b, omega, s)
    for _ in range(nu1 - 1):
        x = jacobi_face(op, x, b, omega, s)
    bc = round_storage(restrict_sum8(b - face_apply(op, x)), s)
    ec = vcycle(h, bc, nu1, nu2, omega, alpha, level + 1)
    xc = x + alpha * prolong_agg(ec, op.shape)   # not stored: fused into the sweep
    so = out_storage or s
    if nu2 >= 1:
        x = jacobi_face(op, xc, b, omega, s if nu2 > 1 else so)
        for q in range(nu2 - 1):
            x = jacobi_face(op, x, b, omega, s if q < nu2 - 2 else so)
    else:
        x = round_storage(xc, so)
    return x


@dataclasses.dataclass
class SolveResult:
    x: np.ndarray
    rel_history: List[float]
    outer_iters: int
    vcycles: int
    converged: bool


def solve_refined(fine: FaceOp, b, tol=1e-8, storage="f32", inner_cycles=4, nu1=2, nu2=2,
                  omega=2.0 / 3.0, alpha=1.0, max_outer=100, hierarchy: Optional[Hierarchy] = None,
                  max_coarse_edge: int = 4) -> SolveResult:
    """FP64 defect correction around low-precision V-cycles.

    Outer step: ``r = b - A x`` in FP64 with the FP32 fine coefficients; stop
    when ``||r||_2 / ||b||_2 <= tol``; otherwise ``e`` = ``inner_cycles``
    V-cycles on ``A e = r`` (r rounded to storage, see ``inner_solve``) and
    ``x += e`` in FP64.
    """
    h = hierarchy or build_hierarchy(fine, storage, max_coarse_edge)
    b64 = np.asarray(b, dtype=np.float64)
    bn = float(np.linalg.norm(b64))
    x = np.zeros_like(b64)
    hist = []
    cycles = 0
    for it in range(max_outer + 1):
        r = b64 - face_apply(h.fine32, x)
        rel = float(np.linalg.norm(r)) / bn if bn > 0 else 0.0
        hist.append(rel)
        if rel <= tol:
            return SolveResult(x, hist, it, cycles, True)
        if it == max_outer or not np.isfinite(rel):
            break
        e = inner_solve(h, r, inner_cycles, nu1, nu2, omega, alpha)
        cycles += inner_cycles
        x = x + e
    return SolveResult(x, hist, len(hist) - 1, cycles, False)


def inner_solve(h: Hierarchy, r, ncycles, nu1=2, nu2=2, omega=2.0 / 3.0, alpha=1.0):
    """``ncycles`` V-cycles on ``A e = r`` in storage precision; the first starts
    from e = 0, later ones use the previous e as the fine-level initial guess."""
    rs = round_storage(r, h.storage)
    e = None
    for _ in range(ncycles):
        e = vcycle(h, rs, nu1, nu2, omega, alpha, 0, e)
    return e


@dataclasses.dataclass
class FCGResult:
    x: np.ndarray
    rel_history: List[float]
    iters: int
    status: str


def _ilogb(v: float) -> int:
    """floor(log2 v) for v > 0 (C's ilogb)."""
    return math.frexp(v)[1] - 1


def fcg_solve(fine: FaceOp, b, tol=1e-8, storage="f32", maxit=200, nu1=2, nu2=2,
              omega=2.0 / 3.0, alpha=1.6, hierarchy: Optional[Hierarchy] = None,
              max_coarse_edge: int = 4, outer: Optional[str] = None,
              f16_rms_exp: int = -4, scaled: Optional[bool] = None) -> FCGResult:
    """FP64 conjugate gradients preconditioned by one storage-precision V-cycle.

    Mirrors the device algorithm step for step (T = storage, P = outer,
    default T):
      r = b, rl = T(r);  per iteration: z = P(V(rl)) from zero;
      rz = rl . z;  beta = rz / rz_prev (0 on the first iteration);
      p = P(fma32(beta, p, z));  q = A p (FP32 coefficients, FP64 arithmetic);
      a = rz / (p . q);  x += a p;  r -= a q;  rl = T(r);
      stop when ||r|| / ||b|| <= tol.
    FP16 storage scales the V-cycle's right-hand side by a power of 2:
    rl = T(2^e r) with e = sbase - ilogb(||r_prev||) (||b|| for the first two),
    sbase = f16_rms_exp + round(log2 sqrt(N)), and rz = (rl . z) / 2^e: FCG is
    invariant to the per-iteration scale of z (p, alpha follow it), which only
    keeps the V-cycle's values inside FP16's range.  ``scaled`` forces the
    scaling on or off for any storage (default: FP16 only).
    """
    h = hierarchy or build_hierarchy(fine, storage, max_coarse_edge)
    s = h.storage
    so = outer or s
    A32 = h.fine32
    b64 = np.asarray(b, dtype=np.float32).astype(np.float64)
    bn2 = float(np.sum(b64 * b64))
    x = np.zeros_like(b64)
    if bn2 == 0.0:
        return FCGResult(x, [0.0], 0, "converged")
    scaled = (s == "f16") if scaled is None else bool(scaled)
    sbase = f16_rms_exp + int(round(0.5 * math.log2(b64.size)))
    sexp = sbase - _ilogb(math.sqrt(bn2)) if scaled else 0
    r = b64.copy()
    rl = round_storage(r * 2.0 ** sexp, s)
    zexp = sexp
    rz = 0.0
    p = None
    hist = [1.0]
    for it in range(maxit):
        z = vcycle(h, rl, nu1, nu2, omega, alpha, out_storage=so)
        rzn = float(np.sum(rl * z)) * 2.0 ** -zexp
        beta = 0.0 if it == 0 else rzn / rz
        rz = rzn
        # one FP32 fused multiply-add, then the storage rounding
        p = round_storage(z if beta == 0.0 else
                          (np.float64(np.float32(beta)) * p + z).astype(np.float32), so)
        q = face_apply(A32, p)
        pap = float(np.sum(p * q))
        if not (pap > 0.0) or not np.isfinite(pap):
            return FCGResult(x, hist, it, "breakdown")
        a = rz / pap
        x = x + a * p
        r = r - a * q
        rl = round_storage(r * 2.0 ** sexp, s)
        rr = float(np.sum(r * r))
        zexp = sexp
        if scaled and rr > 0.0 and np.isfinite(rr):
            sexp = sbase - _ilogb(math.sqrt(rr))
        rel = math.sqrt(rr / bn2)
        hist.append(rel)
        if rel <= tol:
            return FCGResult(x, hist, it + 1, "converged")
        if not np.isfinite(rel):
            return FCGResult(x, hist, it + 1, "breakdown")
    return FCGResult(x, hist, maxit, "maxit")


def solve_fp64_direct(fine: FaceOp, b) -> np.ndarray:
    """FP64 sparse direct solve (the FP64 oracle for small and medium grids)."""
    import scipy.sparse.linalg as spla
    A = face_sparse(fine).tocsc()
    return spla.spsolve(A, np.asarray(b, np.float64).ravel()).reshape(fine.shape)


def solve_fp64_pcg(fine: FaceOp, b, tol=1e-13, maxiter=500) -> np.ndarray:
    """FP64 conjugate gradients preconditioned by one FP64 V(2,2) cycle."""
    h = build_hierarchy(fine, "f64")
    h.ops = [fine.astype("f64")] + [o for o in h.ops[1:]]
    b64 = np.asarray(b, np.float64)
    x = np.zeros_like(b64)
    r = b64.copy()
    z = vcycle(h, r)
    p = z.copy()
    rz = float(np.vdot(r, z))
    bn = float(np.linalg.norm(b64))
    for _ in range(maxiter):
        Ap = face_apply(fine, p)
        a = rz / float(np.vdot(p, Ap))
        x += a * p
        r -= a * Ap
        if np.linalg.norm(r) <= tol * bn:
            break
        z = vcycle(h, r)
        rz_new = float(np.vdot(r, z))
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x


# ---------------------------------------------------------------------------
# Seeded synthetic inputs for the solver configurations
# ---------------------------------------------------------------------------


def make_anisotropic_fields(shape, seed: int, sigma_ln: float = 1.0, n_dipoles: int = 1):
    """Log-normal (mu=0, sigma=sigma_ln) diagonal conductivity and dipole sources.

    Draw order from ``numpy.random.default_rng(seed)``: sx, sy, sz (each a
    (D,H,W) standard-normal field, scaled and exponentiated, stored FP32),
    then for each dipole two distinct interior cells (indices 1..n-2 per
    axis) getting +1 and -1.  Returns (sx, sy, sz, b) as float32 arrays.
    """
    D, H, W = shape
    rng = np.random.default_rng(seed)
    sx = np.exp(sigma_ln * rng.standard_normal(shape)).astype(np.float32)
    sy = np.exp(sigma_ln * rng.standard_normal(shape)).astype(np.float32)
    sz = np.exp(sigma_ln * rng.standard_normal(shape)).astype(np.float32)
    b = np.zeros(shape, np.float32)
    for _ in range(n_dipoles):
        while True:
            p = tuple(int(v) for v in (rng.integers(1, D - 1), rng.integers(1, H - 1), rng.integers(1, W - 1)))
            q = tuple(int(v) for v in (rng.integers(1, D - 1), rng.integers(1, H - 1), rng.integers(1, W - 1)))
            if p != q:
                break
        b[p] += 1.0
        b[q] -= 1.0
    return sx, sy, sz, b


def make_anisotropic_problem(shape, seed: int, sigma_ln: float = 1.0, n_dipoles: int = 1):
    """(FaceOp, b) of ``make_anisotropic_fields`` with harmonic face coefficients."""
    sx, sy, sz, b = make_anisotropic_fields(shape, seed, sigma_ln, n_dipoles)
    return harmonic_faces(sx, sy, sz), b.astype(np.float64)


def batched_instance_specs(n: int, seed: int, edges=(32, 48, 64, 96, 128)):
    """Per-instance (edge, sigma_ln, n_dipoles, seed) for the batched uneven workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        e = int(rng.choice(edges))
        s = float(rng.uniform(0.5, 2.0))
        k = int(rng.integers(1, 5))
        out.append((e, s, k, int(seed * 100003 + i)))
    return out
