"""Reference-shaped operator callables on device tensors.

The reference builds one Keras ``Sequential`` per grid level and operator
(``CNN3D_A_<n>``, ``CNN3D_res_<n>``, ``CNN3D_prol_<n>``,
ART/Anisotropic_3d_Multigrid_7point.ipynb:124-284; ``A_operator_<n>``,
ART/BGS_multigrid_v1.ipynb#c40; the 27-point layers of UM/AI_HFM_v3.ipynb
cells 13-17) and calls them on eager tensors of shape (1, D, H, W, 1)
(or (1, H, W, 1) / (1, W, 1) in 2D / 1D).  Here each is a callable over a
torch tensor on the GPU with the same layout; each call launches a
libhfm.so kernel on torch's current stream and returns a new tensor
(inputs are never mutated, as with the reference's functional layers).

Shape errors raise ValueError (the reference's ``InputLayer`` shape check);
a tensor that is not on a GPU raises ValueError; there is no CPU path.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L

# ---------------------------------------------------------------------------
# tensor plumbing


def _spatial(x: torch.Tensor):
    """(D, H, W) of a (1, ..., 1) tensor with 1-3 spatial axes (or a bare field)."""
    s = tuple(x.shape)
    if len(s) >= 3 and s[0] == 1 and s[-1] == 1:
        s = s[1:-1]
    if not 1 <= len(s) <= 3:
        raise ValueError(f"expected a (1, ..., 1) field with 1-3 spatial axes, got {tuple(x.shape)}")
    return (1,) * (3 - len(s)) + s, len(s)


def _prep(x, name="x") -> torch.Tensor:
    if not isinstance(x, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on the GPU")
    if x.device.type != "cuda":
        raise ValueError(f"{name} is on {x.device}; libhfm kernels need a GPU tensor")
    if x.dtype != torch.float32:
        x = x.to(torch.float32)  # the reference's layers cast float64 inputs to float32
    return x.contiguous()


def kernel27(K) -> np.ndarray:
    """Reference weight array (1,3,3,3,1) / (3,3,3) / (1,3,3,1) / (3,3) / (1,3,1) / (3,)
    as the 27 row-major taps of a 3x3x3 kernel (2D/1D kernels sit in the middle planes)."""
    K = np.asarray(K, dtype=np.float64)
    K = np.squeeze(K)
    if K.shape == (3, 3, 3):
        K3 = K
    elif K.shape == (3, 3):
        K3 = np.zeros((3, 3, 3)); K3[1] = K
    elif K.shape == (3,):
        K3 = np.zeros((3, 3, 3)); K3[1, 1] = K
    else:
        raise ValueError(f"unsupported stencil weight shape {np.shape(K)}")
    return np.ascontiguousarray(K3.astype(np.float32).ravel())


def _kptr(K27: np.ndarray):
    return K27.ctypes.data_as(L.C.c_void_p)


# ---------------------------------------------------------------------------
# stencil layers


class FixedConv:
    """``ConvNd(1, kernel_size=3, strides=1, padding='SAME')`` with fixed weights, bias 0.

    ``input_shape`` (spatial, optional) mirrors the reference's ``InputLayer``:
    calls with another shape raise ValueError.
    """

    def __init__(self, weights, input_shape: Optional[Sequence[int]] = None):
        self.K = kernel27(weights)
        self.input_shape = tuple(input_shape) if input_shape is not None else None
        self.center = float(self.K[13])

    def _check(self, x):
        dims, _ = _spatial(x)
        if self.input_shape is not None:
            want = (1,) * (3 - len(self.input_shape)) + self.input_shape
            if dims != want:
                raise ValueError(f"layer expects spatial shape {self.input_shape}, got {tuple(x.shape)}")
        return dims

    def __call__(self, x):
        x = _prep(x)
        D, H, W = self._check(x)
        y = torch.empty_like(x)
        L.call("hfm_conv27_f32", x.data_ptr(), y.data_ptr(), 0, _kptr(self.K), 1.0, D, H, W,
               L.FORM["apply"], L.stream_handle(x.device))
        return y

    def residual(self, x, b, zero_outlet: bool = False):
        """``A(x) - b`` in one pass (the reference's ``r = CNN3D_A(values) - b``)."""
        return self._form(x, b, "resid", zero_outlet)

    def residual_restricted(self, x, b, zero_outlet: bool = False):
        """``restrict_mean(A(x) - b)`` in one pass, without the fine residual
        (UM/AI_HFM_v3.ipynb:2573-2575: ``r = CNN3D_A_512(values_p) - b``, the
        outlet plane zeroed, ``r_256 = CNN3D_res_512(r)``); bitwise the
        ``residual`` + ``restrict_mean`` pair.  3D fields with even extents."""
        x, b = _prep(x), _prep(b, "b")
        D, H, W = self._check(x)
        (_, _, _), nd = _spatial(x)
        if b.shape != x.shape:
            raise ValueError("b must have the shape of x")
        if nd != 3 or D % 2 or H % 2 or W % 2:
            raise ValueError("residual_restricted needs a 3D field with even extents")
        out_shape = list(x.shape)
        off = 1 if x.dim() > nd else 0
        for a in range(nd):
            out_shape[off + a] //= 2
        y = torch.empty(out_shape, dtype=torch.float32, device=x.device)
        L.call("hfm_resid_restrict_f32", x.data_ptr(), b.data_ptr(), y.data_ptr(), _kptr(self.K),
               D, H, W, L.FORM_ZERO_OUTLET if zero_outlet else 0, L.stream_handle(x.device))
        return y

    def subtract_from(self, b, x):
        """``b - A(x)`` in one pass (the reference's ``values_u - CNN3D_pu(values_p)``)."""
        return self._form(x, b, "subtract", False)

    def axpy(self, b, x, scale: float):
        """``b + scale * A(x)`` in one pass (explicit stage updates)."""
        return self._form(x, b, "axpy", False, scale)

    def _form(self, x, b, form, zero_outlet, d=1.0):
        x, b = _prep(x), _prep(b, "b")
        D, H, W = self._check(x)
        if b.shape != x.shape:
            raise ValueError("b must have the shape of x")
        y = torch.empty_like(x)
        f = L.FORM[form] | (L.FORM_ZERO_OUTLET if zero_outlet else 0)
        L.call("hfm_conv27_f32", x.data_ptr(), y.data_ptr(), b.data_ptr(), _kptr(self.K), float(d),
               D, H, W, f, L.stream_handle(x.device))
        return y

    def jacobi(self, w, r, d: Optional[float] = None):
        """``w - A(w)/d + r/d`` in one pass; ``d`` defaults to the kernel centre."""
        w, r = _prep(w, "w"), _prep(r, "r")
        D, H, W = self._check(w)
        if r.shape != w.shape:
            raise ValueError("r must have the shape of w")
        dd = self.center if d is None else float(np.float32(d))
        y = torch.empty_like(w)
        L.call("hfm_conv27_f32", w.data_ptr(), y.data_ptr(), r.data_ptr(), _kptr(self.K), dd,
               D, H, W, L.FORM["jacobi"], L.stream_handle(w.device))
        return y


VAR6_DIRS = ("xm1", "xp1", "ym1", "yp1", "zm1", "zp1")


class AOperatorVar6:
    """``A_operator_n`` of ART/BGS_multigrid_v1.ipynb#c40: sum over the six
    directions of ``A_MAT_dir * (x - x_nb)``; ``mats`` maps xm1..zp1 (and
    optionally ``tot``) to (1, D, H, W, 1) coefficient tensors."""

    def __init__(self, mats: Dict[str, torch.Tensor]):
        self.m = [_prep(mats[k], k) for k in VAR6_DIRS]
        tot = mats.get("tot")
        if tot is None:
            tot = sum(self.m)
        self.tot = _prep(tot, "tot")
        self.dims, _ = _spatial(self.m[0])
        for t in (*self.m, self.tot):
            if _spatial(t)[0] != self.dims:
                raise ValueError("all coefficient fields must share one shape")
        self._arr = L.ptr_array([t.data_ptr() for t in self.m])

    def _run(self, x, b, form):
        x = _prep(x)
        dims, _ = _spatial(x)
        if dims != self.dims:
            raise ValueError(f"operator built for {self.dims}, got {dims}")
        y = torch.empty_like(x)
        L.call("hfm_var6_f32", x.data_ptr(), y.data_ptr(), L.ptr(b), self._arr,
               self.tot.data_ptr(), *dims, form, L.stream_handle(x.device))
        return y

    def __call__(self, x):
        return self._run(x, None, L.FORM["apply"])

    def residual(self, x, b):
        return self._run(x, _prep(b, "b"), L.FORM["resid"])

    def jacobi(self, w, r):
        """``w - A(w)/A_MAT_TOT + r/A_MAT_TOT`` (BGS ``#c42``)."""
        return self._run(w, _prep(r, "r"), L.FORM["jacobi"])


def mesh_shift(mesh: np.ndarray, axis: int) -> np.ndarray:
    """``mesh_*_p1`` of ART/BGS_multigrid_v1.ipynb (cell 36): the grid-size
    array shifted by one cell toward lower index along ``axis`` on the
    interior (``p1[i] = mesh[i + 1]`` for 1 <= i <= n - 2), both boundary
    cells keeping their own size.  ``mesh`` is the (D, H, W) or
    (1, D, H, W, 1) array; axis counts over (D, H, W)."""
    m = np.asarray(mesh)
    grid = m[0, ..., 0] if m.ndim == 5 else m
    out = grid.copy()
    n = grid.shape[axis]
    if n > 2:
        dst = [slice(None)] * 3
        src = [slice(None)] * 3
        dst[axis] = slice(1, n - 1)
        src[axis] = slice(2, n)
        out[tuple(dst)] = grid[tuple(src)]
    return out.reshape(m.shape)


def nonuniform_coefficients(mesh_x, mesh_y, mesh_z, Dx=100.0, Dy=100.0, Dz=100.0) -> Dict:
    """The fine-level ``A_MAT_*`` arrays of ART/BGS_multigrid_v1.ipynb (cell 38)
    on a non-uniform mesh: ``A_MAT_xm1 = Dx / mesh_x^2``, ``A_MAT_xp1 = Dx /
    mesh_x_p1^2`` (likewise y, z) and ``A_MAT_TOT`` their sum, in float32
    (host arrays of the field's shape; scalars or arrays for Dx, Dy, Dz)."""
    f32 = np.float32
    shape = np.shape(mesh_x)
    mats = {}
    for name, mesh, D, axis in (("x", mesh_x, Dx, 2), ("y", mesh_y, Dy, 1), ("z", mesh_z, Dz, 0)):
        mesh = np.asarray(mesh, dtype=np.float64)
        if mesh.shape != shape:
            raise ValueError("mesh_x, mesh_y and mesh_z must share one shape")
        D = np.broadcast_to(np.asarray(D, dtype=np.float64), shape)
        mats[name + "m1"] = (D / mesh ** 2).astype(f32)
        mats[name + "p1"] = (D / mesh_shift(mesh, axis) ** 2).astype(f32)
    tot = mats["xm1"] + mats["ym1"]
    for name in ("zm1", "xp1", "yp1", "zp1"):
        tot = tot + mats[name]
    mats["tot"] = tot
    return mats


def var6_hierarchy(mats: Dict, nlevels: int, device=None) -> list:
    """``AOperatorVar6`` per level from fine-level coefficient arrays: every
    coarser level's fields (TOT included) are the 2^d mean restriction of the
    finer level's (ART/BGS_multigrid_v1.ipynb cell 38, ``CNN3D_res_n`` of each
    ``A_MAT_*_n``), computed on the GPU."""
    dev = torch.device(device or "cuda")
    level = {k: torch.as_tensor(np.asarray(v, np.float32)).to(dev) for k, v in mats.items()}
    if "tot" not in level:
        raise ValueError("mats needs the 'tot' array (nonuniform_coefficients makes it)")
    ops = [AOperatorVar6(level)]
    for _ in range(1, nlevels):
        level = {k: restrict_mean(v) for k, v in level.items()}
        ops.append(AOperatorVar6(level))
    return ops


def _factors(nd):
    return (1,) * (3 - nd) + (2,) * nd


def restrict_mean(x):
    """``CNN3D_res_n``: stride-2 mean of the 2^d children (…7point.ipynb:194-235)."""
    x = _prep(x)
    (D, H, W), nd = _spatial(x)
    fz, fy, fx = _factors(nd)
    out_shape = list(x.shape)
    off = 1 if x.dim() > nd else 0
    for a in range(nd):
        out_shape[off + a] //= 2
    y = torch.empty(out_shape, dtype=torch.float32, device=x.device)
    L.call("hfm_restrict_mean_f32", x.data_ptr(), y.data_ptr(), D, H, W, fz, fy, fx,
           L.stream_handle(x.device))
    return y


def prolong_nearest(x):
    """``CNN3D_prol_n``: UpSampling by 2 on every spatial axis (…7point.ipynb:251-284)."""
    x = _prep(x)
    (D, H, W), nd = _spatial(x)
    fz, fy, fx = _factors(nd)
    out_shape = list(x.shape)
    off = 1 if x.dim() > nd else 0
    for a in range(nd):
        out_shape[off + a] *= 2
    y = torch.empty(out_shape, dtype=torch.float32, device=x.device)
    L.call("hfm_prolong_nearest_f32", x.data_ptr(), y.data_ptr(), D, H, W, fz, fy, fx,
           L.stream_handle(x.device))
    return y


# ---------------------------------------------------------------------------
# the reference's sawtooth cycle


class SawtoothMG:
    """The reference's multigrid cycle (…7point.ipynb:402-454, BGS ``#c42``).

    ``ops[l]`` is a FixedConv (constant weights, divisor = its centre or
    ``divisors[l]``) or an AOperatorVar6 for level l (fine first).  One call
    of ``cycle`` does: r = A x - b; restrict r to every level; from w = 0 do
    ``coarsest_sweeps`` sweeps (default ``j_itr``) on the coarsest level; at
    each finer level prolong and do ``j_itr`` sweeps against that level's r;
    ``x -= w``; one fine sweep on A x = b; optional boundary fill ``bc``.

    ``fine_sweep=False`` is the urban drivers' form: the finest level takes
    the prolonged coarse correction with no sweep of its own
    (UM/AI_HFM_v3.ipynb:2610-2611, UM/INHALE_demo.ipynb:13454-13458,
    ARTs/naiver_stokes_3D.py:696-699).  Prolongation + first sweep,
    correction + fine sweep and (without the fine-level sweep) prolongation +
    correction + fine sweep are single fused kernels.
    """

    def __init__(self, ops: List, divisors: Optional[List[float]] = None,
                 coarsest_sweeps: Optional[int] = None, bc: Optional[Callable] = None,
                 zero_outlet: bool = False, fine_sweep: bool = True, j_itr: int = 1):
        self.ops = ops
        self.divisors = divisors
        self.j_itr = int(j_itr)
        if self.j_itr < 1:
            raise ValueError("j_itr must be at least 1")
        self.coarsest_sweeps = self.j_itr if coarsest_sweeps is None else int(coarsest_sweeps)
        self.bc = bc
        self.zero_outlet = bool(zero_outlet)
        self.fine_sweep = bool(fine_sweep)

    def _d(self, lev):
        op = self.ops[lev]
        if isinstance(op, AOperatorVar6):
            return None
        if self.divisors is not None:
            return float(np.float32(self.divisors[lev]))
        return op.center

    def _sweep(self, lev, w, r):
        op = self.ops[lev]
        return op.jacobi(w, r) if isinstance(op, AOperatorVar6) else op.jacobi(w, r, self._d(lev))

    def _op_args(self, lev):
        op = self.ops[lev]
        if isinstance(op, AOperatorVar6):
            return None, 0.0, op._arr, op.tot.data_ptr()
        return _kptr(op.K), self._d(lev), None, None

    def _prolong_jacobi(self, lev, wc, r):
        r = _prep(r, "r")
        (D, H, W), nd = _spatial(r)
        fz, fy, fx = _factors(nd)
        if wc.numel() * fz * fy * fx != D * H * W:
            raise ValueError(f"coarse correction of {wc.numel()} cells does not match level {(D, H, W)}")
        y = torch.empty_like(r)
        K, d, m, tot = self._op_args(lev)
        L.call("hfm_prolong_jacobi_f32", wc.data_ptr(), r.data_ptr(), y.data_ptr(), K, d, m, tot,
               D, H, W, fz, fy, fx, L.stream_handle(r.device))
        return y

    def _correct_jacobi(self, x, w, b, prolonged: bool):
        """J(x - w) against b; with ``prolonged`` w is the next level's
        correction, applied through the nearest prolongation in the kernel."""
        flags = L.CORRECT_ZERO_OUTLET if self.zero_outlet else 0
        (D, H, W), nd = _spatial(x)
        if b.shape != x.shape:
            raise ValueError("x and b must share a shape")
        y = torch.empty_like(x)
        s = L.stream_handle(x.device)
        K, d, m, tot = self._op_args(0)
        if prolonged:
            fz, fy, fx = _factors(nd)
            if w.numel() * fz * fy * fx != D * H * W:
                raise ValueError(f"coarse correction of {w.numel()} cells does not match {(D, H, W)}")
            L.call("hfm_prolong_correct_jacobi_f32", x.data_ptr(), w.data_ptr(), b.data_ptr(),
                   y.data_ptr(), K, d, m, tot, D, H, W, fz, fy, fx, flags, s)
        else:
            if w.shape != x.shape:
                raise ValueError("x, w and b must share a shape")
            L.call("hfm_correct_jacobi_f32", x.data_ptr(), w.data_ptr(), b.data_ptr(), y.data_ptr(),
                   K, d, m, tot, D, H, W, flags, s)
        return y

    def cycle(self, values, b, record: Optional[dict] = None):
        """One cycle; with ``zero_outlet`` the fine residual and the corrected
        field are zeroed on the x = W-1 plane (UM/AI_HFM_v3.ipynb:2574,2611-2612).
        ``record[l]`` receives level l's correction; without the fine-level
        sweep the finest correction (the prolonged ``record[1]``) is never
        formed, and ``record[0]`` is left unset."""
        values, b = _prep(values, "values"), _prep(b, "b")
        nlev = len(self.ops)
        (D, H, W), nd = _spatial(values)
        if (not self.fine_sweep and nlev > 1 and isinstance(self.ops[0], FixedConv) and nd == 3
                and D % 2 == 0 and H % 2 == 0 and W % 2 == 0):
            # the finest level takes no sweep: its residual has no reader but
            # the restriction, and the two run as one pass
            rs = [None, self.ops[0].residual_restricted(values, b, zero_outlet=self.zero_outlet)]
        else:
            r = self.ops[0].residual(values, b, zero_outlet=self.zero_outlet) if isinstance(
                self.ops[0], FixedConv) else self.ops[0].residual(values, b)
            rs = [r]
        while len(rs) < nlev:
            rs.append(restrict_mean(rs[-1]))
        w = torch.zeros_like(rs[-1])
        for _ in range(self.coarsest_sweeps):
            w = self._sweep(nlev - 1, w, rs[-1])
        if record is not None:
            record[nlev - 1] = w
        low = 0 if (self.fine_sweep or nlev == 1) else 1
        for lev in range(nlev - 2, low - 1, -1):
            w = self._prolong_jacobi(lev, w, rs[lev])
            for _ in range(self.j_itr - 1):
                w = self._sweep(lev, w, rs[lev])
            if record is not None:
                record[lev] = w
        values = self._correct_jacobi(values, w, b, prolonged=low == 1)
        if self.bc is not None:
            values = self.bc(values)
        return values


def prolong_upsum(x):
    """``UpSampling2D(2)`` + ``Conv2D(2x2 of ones, padding='SAME')`` of
    ARTn/Multigrid_cnn.ipynb:342-398 (any 1-3 spatial axes): output i sums the
    upsampled cells i and i + 1 of every axis, zero past the end."""
    x = _prep(x)
    (D, H, W), nd = _spatial(x)
    fz, fy, fx = _factors(nd)
    out_shape = list(x.shape)
    off = 1 if x.dim() > nd else 0
    for a in range(nd):
        out_shape[off + a] *= 2
    y = torch.empty(out_shape, dtype=torch.float32, device=x.device)
    L.call("hfm_prolong_upsum_f32", x.data_ptr(), y.data_ptr(), D, H, W, fz, fy, fx,
           L.stream_handle(x.device))
    return y


def field_sum(x) -> torch.Tensor:
    """Sum of a field's values in FP64 (fixed order) as a one-element device tensor."""
    x = _prep(x)
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    L.call("hfm_sum_f64", x.data_ptr(), x.numel(), out.data_ptr(), L.stream_handle(x.device))
    return out


class MultigridCNN:
    """The 2D multigrid of ARTn/Multigrid_cnn.ipynb (cells 2-21).

    One fixed 3x3 operator ``A`` at every level (``w3``), a smoother whose
    divisor is a layer weight, ``w - bias(A(w)) + bias(r)`` with ``bias`` the
    centre weight ``c`` (1/107), 2x2 mean restriction down to 1 x 1, and the
    upsampling prolongation ``prolong_upsum``.  A cycle: pre-sweep; r = A x - b;
    restrict; from w = 0, ``j_itr`` sweeps at the coarsest level and after each
    prolongation below the finest; at the finest, the prolongation and one
    sweep against r; x -= w; post-sweep.  ``last_norm`` is the notebook's
    residual monitor ``sum(r) / (nx ny)`` of the last cycle (a device tensor).
    """

    def __init__(self, weights=None, bias: float = 1.0 / 107.0, j_itr: int = 2):
        K = np.array([[0.0, -1.5, 0.0], [-51.5, 107.0, -1.5], [0.0, -51.5, 0.0]]) \
            if weights is None else weights
        self.A = FixedConv(K)
        self.c = float(np.float32(bias))
        self.j_itr = int(j_itr)
        self._rsum, self._n = None, 0

    def _sweep(self, w, r):
        return self.A._form(w, r, "jacobi_mul", False, self.c)

    @property
    def last_norm(self) -> Optional[float]:
        """sum(r) / (nx ny) of the last cycle's residual (reads it back)."""
        return None if self._rsum is None else float(self._rsum) / self._n

    def cycle(self, values, b):
        values, b = _prep(values, "values"), _prep(b, "b")
        if b.shape != values.shape:
            raise ValueError("values and b must share a shape")
        values = self._sweep(values, b)
        r = self.A.residual(values, b)
        rs = [r]
        while True:
            dims, nd = _spatial(rs[-1])
            if min(dims[3 - nd:]) <= 1:
                break
            rs.append(restrict_mean(rs[-1]))
        w = torch.zeros_like(rs[-1])
        for _ in range(self.j_itr):
            w = self._sweep(w, rs[-1])
        for lev in range(len(rs) - 2, 0, -1):
            w = prolong_upsum(w)
            for _ in range(self.j_itr):
                w = self._sweep(w, rs[lev])
        if len(rs) > 1:
            w = self._sweep(prolong_upsum(w), r)
        # values - w and the post-sweep in one pass
        (D, H, W), _ = _spatial(values)
        out = torch.empty_like(values)
        L.call("hfm_correct_jacobi_f32", values.data_ptr(), w.data_ptr(), b.data_ptr(),
               out.data_ptr(), _kptr(self.A.K), self.c, None, None, D, H, W,
               L.CORRECT_RECIPROCAL, L.stream_handle(values.device))
        self._rsum, self._n = field_sum(r), r.numel()
        return out


def level_layers(K, fine_shape: Sequence[int], nlevels: int, prefix: str = "CNN3D") -> dict:
    """The reference's per-level layer registry (UM/AI_HFM_v3.ipynb:587-623):
    ``{prefix}_A_<tag>``, ``{prefix}_res_<tag>``, ``{prefix}_prol_<tag>`` with
    tag = 2^(nlevels-1-l) for level l (fine first), as built with ``locals()``."""
    out = {}
    shape = tuple(fine_shape)
    for lev in range(nlevels):
        tag = 2 ** (nlevels - 1 - lev)
        s = tuple(n // 2 ** lev for n in shape)
        out[f"{prefix}_A_{tag}"] = FixedConv(K, s)
        if lev < nlevels - 1:
            out[f"{prefix}_res_{tag}"] = restrict_mean
        if lev > 0:
            out[f"{prefix}_prol_{tag}"] = prolong_nearest
    return out


# ---------------------------------------------------------------------------
# boundary and obstacle functions (UM/AI_HFM_v3.ipynb cell 19)


def _bc(field, kind):
    f = _prep(field)
    (D, H, W), _ = _spatial(f)
    out = torch.empty_like(f)
    L.call("hfm_bc_f32", f.data_ptr(), out.data_ptr(), D, H, W, L.BC[kind],
           L.stream_handle(f.device))
    return out


def _check_nx(f, nx):
    if nx is not None and _spatial(f)[0][2] != int(nx):
        raise ValueError(f"nx={nx} does not match the field's x extent {_spatial(f)[0][2]}")


def boundary_condition_top(values):
    """Neumann copy of plane 1 into plane 0 of axis 1 (…_BC.ipynb:340-344)."""
    return _bc(values, "top_copy")


def boundary_condition_pressure(values_p, nx=None):
    _check_nx(values_p, nx)
    return _bc(values_p, "pressure")


def boundary_condition_source(b, nx=None):
    _check_nx(b, nx)
    return _bc(b, "source")


def boundary_condition_indicator(alpha, nx=None):
    _check_nx(alpha, nx)
    return _bc(alpha, "indicator")


def zero_outlet(r):
    """``r[..., nx-1] = 0`` (UM/AI_HFM_v3.ipynb:2574,2612)."""
    return _bc(r, "zero_outlet")


def boundary_condition_velocity(values_u, values_v, values_w, nx=None, ub: float = 1.0,
                                inflow=None):
    """Velocity boundary planes; ``inflow`` (D values, optional) replaces the constant
    ``ub`` on the x faces by a height profile."""
    u, v, w = (_prep(a, n) for a, n in ((values_u, "u"), (values_v, "v"), (values_w, "w")))
    _check_nx(u, nx)
    (D, H, W), _ = _spatial(u)
    if v.shape != u.shape or w.shape != u.shape:
        raise ValueError("u, v, w must share a shape")
    if inflow is not None:
        inflow = _prep(inflow, "inflow")
        if inflow.numel() != D:
            raise ValueError("inflow profile needs one value per height")
    uo, vo, wo = torch.empty_like(u), torch.empty_like(v), torch.empty_like(w)
    L.call("hfm_bc_velocity_f32", u.data_ptr(), v.data_ptr(), w.data_ptr(), uo.data_ptr(),
           vo.data_ptr(), wo.data_ptr(), D, H, W, float(ub), L.ptr(inflow),
           L.stream_handle(u.device))
    return uo, vo, wo


def boundary_condition_velocity_(u, v, w, ub: float = 1.0, inflow=None, scratch=None):
    """``boundary_condition_velocity`` in place: writes only the boundary planes
    of u, v, w (bitwise the out-of-place result) and returns them.  For a
    driver's own intermediates: the reference's function returns new tensors.
    ``scratch``: 2 D W + 2 H W floats on the device (allocated when omitted)."""
    for a, n in ((u, "u"), (v, "v"), (w, "w")):
        if not (isinstance(a, torch.Tensor) and a.device.type == "cuda"
                and a.dtype == torch.float32 and a.is_contiguous()):
            raise ValueError(f"{n} must be a contiguous float32 GPU tensor to be filled in place")
    (D, H, W), _ = _spatial(u)
    if v.shape != u.shape or w.shape != u.shape:
        raise ValueError("u, v, w must share a shape")
    if inflow is not None:
        inflow = _prep(inflow, "inflow")
        if inflow.numel() != D:
            raise ValueError("inflow profile needs one value per height")
    need = 2 * D * W + 2 * H * W
    if scratch is None:
        scratch = torch.empty(need, dtype=torch.float32, device=u.device)
    elif scratch.numel() < need or scratch.dtype != torch.float32 or not scratch.is_contiguous():
        raise ValueError(f"scratch needs {need} contiguous float32 values")
    L.call("hfm_bc_velocity_inplace_f32", u.data_ptr(), v.data_ptr(), w.data_ptr(),
           scratch.data_ptr(), D, H, W, float(ub), L.ptr(inflow), L.stream_handle(u.device))
    return u, v, w


def boundary_condition_velocity_bluff_(u, v, w, sigma, dt: float, ub: float = 1.0, inflow=None,
                                       scratch=None):
    """``bluff_body(*boundary_condition_velocity_(u, v, w))`` in place, after
    ``gradient3(..., post_bluff=True)`` (which leaves the cells the fill copies
    from undamped): bitwise the plain gradient, the in-place fill and the
    damping as three passes."""
    for a, n in ((u, "u"), (v, "v"), (w, "w")):
        if not (isinstance(a, torch.Tensor) and a.device.type == "cuda"
                and a.dtype == torch.float32 and a.is_contiguous()):
            raise ValueError(f"{n} must be a contiguous float32 GPU tensor to be filled in place")
    (D, H, W), _ = _spatial(u)
    sigma = _prep(sigma, "sigma")
    if v.shape != u.shape or w.shape != u.shape or sigma.numel() != u.numel():
        raise ValueError("u, v, w and sigma must share a shape")
    if inflow is not None:
        inflow = _prep(inflow, "inflow")
        if inflow.numel() != D:
            raise ValueError("inflow profile needs one value per height")
    need = 2 * D * W + 2 * H * W
    if scratch is None:
        scratch = torch.empty(need, dtype=torch.float32, device=u.device)
    elif scratch.numel() < need or scratch.dtype != torch.float32 or not scratch.is_contiguous():
        raise ValueError(f"scratch needs {need} contiguous float32 values")
    L.call("hfm_bc_velocity_bluff_inplace_f32", u.data_ptr(), v.data_ptr(), w.data_ptr(),
           scratch.data_ptr(), D, H, W, float(ub), L.ptr(inflow), sigma.data_ptr(), float(dt),
           L.stream_handle(u.device))
    return u, v, w


def bluff_body(values_u, values_v, values_w, alpha, sigma, dt: float):
    """``f / (1 + dt*sigma)`` for u, v, w and the scalar (immersed obstacle
    damping); ``alpha=None`` damps the velocities only (the INHALE demo's
    ``bluff_body(u, v, w, sigma)``) and returns three fields."""
    fields = [_prep(a) for a in (values_u, values_v, values_w, alpha) if a is not None]
    sigma = _prep(sigma, "sigma")
    for f in fields:
        if f.shape != sigma.shape:
            raise ValueError("fields and sigma must share a shape")
    outs = [torch.empty_like(f) for f in fields]
    L.call("hfm_bluff_body_f32", L.ptr_array([f.data_ptr() for f in fields]),
           L.ptr_array([o.data_ptr() for o in outs]), len(fields), sigma.data_ptr(),
           float(dt), sigma.numel(), L.stream_handle(sigma.device))
    return tuple(outs)


# ---------------------------------------------------------------------------
# fused time-step kernels (UM/AI_HFM_v3.ipynb cell 25)


def advect(vel, phis, bases, f: float, kernels, indicator: bool = False, sigma=None,
           dt: float = 0.0, nbluff: int = 0):
    """``out_q = f * (dif(phi_q) - u ax(phi_q) - v ay(phi_q) - w az(phi_q)) + base_q``
    for 1 or 3 fields in one pass; ``kernels`` = (dif, ax, ay, az) weight arrays.
    ``indicator``: then ``boundary_condition_indicator``; ``nbluff`` (0-2): then
    that many ``bluff_body`` dampings with ``sigma``, ``dt`` -- in the same pass,
    bitwise the separate calls."""
    post = (L.ADV_INDICATOR if indicator else 0) | (int(nbluff) << L.ADV_BLUFF_SHIFT)
    if not 0 <= int(nbluff) <= 2:
        raise ValueError("nbluff must be 0, 1 or 2")
    if nbluff and sigma is None:
        raise ValueError("damping needs sigma")
    vel = [_prep(a, "vel") for a in vel]
    phis = [_prep(a, "phi") for a in phis]
    bases = [_prep(a, "base") for a in bases]
    if len(vel) != 3 or len(phis) not in (1, 3) or len(bases) != len(phis):
        raise ValueError("advect takes 3 velocity fields and 1 or 3 fields with bases")
    shape = vel[0].shape
    for t in (*vel, *phis, *bases):
        if t.shape != shape:
            raise ValueError("all fields must share one shape")
    (D, H, W), _ = _spatial(vel[0])
    outs = [torch.empty_like(a) for a in phis]
    K = np.ascontiguousarray(np.stack([kernel27(k) for k in kernels]), dtype=np.float32)
    if post:
        sig = _prep(sigma, "sigma") if sigma is not None else None
        if sig is not None and sig.numel() != vel[0].numel():
            raise ValueError("sigma must match the fields")
        L.call("hfm_advect27_post_f32", L.ptr_array([a.data_ptr() for a in vel]),
               L.ptr_array([a.data_ptr() for a in phis]),
               L.ptr_array([a.data_ptr() for a in bases]),
               L.ptr_array([a.data_ptr() for a in outs]), len(phis),
               K.ctypes.data_as(L.C.c_void_p), float(f), L.ptr(sig), float(dt), post, D, H, W,
               L.stream_handle(vel[0].device))
        return outs
    L.call("hfm_advect27_f32", L.ptr_array([a.data_ptr() for a in vel]),
           L.ptr_array([a.data_ptr() for a in phis]), L.ptr_array([a.data_ptr() for a in bases]),
           L.ptr_array([a.data_ptr() for a in outs]), len(phis), K.ctypes.data_as(L.C.c_void_p),
           float(f), D, H, W, L.stream_handle(vel[0].device))
    return outs


def gradient3(p, fields, kernels, sigma=None, dt: float = 0.0, p_bc: bool = False,
              pre_bluff: bool = False, post_bluff: bool = False, out=None):
    """``f_q - K_q (*) p`` for the three velocity components in one walk over p
    (the urban step's ``values_u - CNN3D_Su(values_p)``, ...).  ``p_bc`` reads p
    through ``boundary_condition_pressure`` and also returns the filled p;
    ``pre_bluff`` damps f first (``bluff_body`` before the gradient);
    ``post_bluff`` damps the result except where the velocity fill copies from
    (finish with ``boundary_condition_velocity_bluff_``).  ``out``: output
    tensors (may be the fields themselves).  Returns (u, v, w, p_filled or None)."""
    p = _prep(p, "p")
    fields = [_prep(a, "f") for a in fields]
    if len(fields) != 3 or any(f.shape != p.shape for f in fields):
        raise ValueError("gradient3 takes p and three fields of one shape")
    if (pre_bluff or post_bluff) and sigma is None:
        raise ValueError("damping needs sigma")
    sig = _prep(sigma, "sigma") if sigma is not None else None
    if sig is not None and sig.numel() != p.numel():
        raise ValueError("sigma must match the fields")
    (D, H, W), _ = _spatial(p)
    outs = [torch.empty_like(f) for f in fields] if out is None else list(out)
    pout = torch.empty_like(p) if p_bc else None
    flags = ((L.GRAD_P_BC if p_bc else 0) | (L.GRAD_PRE_BLUFF if pre_bluff else 0)
             | (L.GRAD_POST_BLUFF if post_bluff else 0))
    K = np.ascontiguousarray(np.stack([kernel27(k) for k in kernels]), dtype=np.float32)
    L.call("hfm_grad3_f32", p.data_ptr(), L.ptr_array([f.data_ptr() for f in fields]),
           L.ptr_array([o.data_ptr() for o in outs]), L.ptr(pout),
           K.ctypes.data_as(L.C.c_void_p), L.ptr(sig), float(dt), flags, D, H, W,
           L.stream_handle(p.device))
    return outs[0], outs[1], outs[2], pout


def advect_dissipative(vel, phis, bases, dis, f: float, kernels):
    """``out_q = f * (dis_q - u ax(phi_q) - v ay(phi_q) - w az(phi_q)) + base_q``:
    the momentum stage of ``UM/INHALE_demo.ipynb`` cell 24 with a given
    dissipation field (``petrov_galerkin_dissipation``) in place of the
    diffusion layer; ``kernels`` = (ax, ay, az) weight arrays."""
    vel = [_prep(a, "vel") for a in vel]
    phis = [_prep(a, "phi") for a in phis]
    bases = [_prep(a, "base") for a in bases]
    dis = [_prep(a, "dis") for a in dis]
    if len(vel) != 3 or len(phis) not in (1, 3) or len(bases) != len(phis) or len(dis) != len(phis):
        raise ValueError("advect_dissipative takes 3 velocity fields and 1 or 3 fields with "
                         "bases and dissipations")
    shape = vel[0].shape
    for t in (*vel, *phis, *bases, *dis):
        if t.shape != shape:
            raise ValueError("all fields must share one shape")
    (D, H, W), _ = _spatial(vel[0])
    outs = [torch.empty_like(a) for a in phis]
    K = np.ascontiguousarray(np.stack([kernel27(k) for k in kernels]), dtype=np.float32)
    if K.shape != (3, 27):
        raise ValueError("kernels = (ax, ay, az)")
    L.call("hfm_advect27_dis_f32", L.ptr_array([a.data_ptr() for a in vel]),
           L.ptr_array([a.data_ptr() for a in phis]), L.ptr_array([a.data_ptr() for a in bases]),
           L.ptr_array([a.data_ptr() for a in dis]), L.ptr_array([a.data_ptr() for a in outs]),
           len(phis), K.ctypes.data_as(L.C.c_void_p), float(f), D, H, W,
           L.stream_handle(vel[0].device))
    return outs


def petrov_galerkin_dissipation(values_u, values_v, values_w, eplsion_k, sigma, dt: float,
                                kernels):
    """``Petrov_Galerkin_dissipation`` of ``UM/INHALE_demo.ipynb`` (cell 20):
    returns (k_x, k_y, k_z).  ``kernels`` = (dif, xadv, yadv, zadv) weight arrays
    (CNN3D_central_2nd_*).  Three launches: the damped velocities
    f / (1 + dt sigma), the diffusivities k (one neighbourhood load per
    component for the four stencils), the flux."""
    u, v, w = (_prep(a, n) for a, n in ((values_u, "u"), (values_v, "v"), (values_w, "w")))
    sigma = _prep(sigma, "sigma")
    for t in (v, w, sigma):
        if t.shape != u.shape:
            raise ValueError("u, v, w and sigma must share a shape")
    (D, H, W), _ = _spatial(u)
    s = L.stream_handle(u.device)
    vt = [torch.empty_like(u) for _ in range(3)]
    L.call("hfm_bluff_body_f32", L.ptr_array([a.data_ptr() for a in (u, v, w)]),
           L.ptr_array([a.data_ptr() for a in vt]), 3, sigma.data_ptr(), float(dt), u.numel(), s)
    K = np.ascontiguousarray(np.stack([kernel27(k) for k in kernels]), dtype=np.float32)
    if K.shape != (4, 27):
        raise ValueError("kernels = (dif, xadv, yadv, zadv)")
    ks = [torch.empty_like(u) for _ in range(3)]
    L.call("hfm_pg_coef_f32", L.ptr_array([a.data_ptr() for a in vt]), sigma.data_ptr(),
           L.ptr_array([a.data_ptr() for a in ks]), K.ctypes.data_as(L.C.c_void_p), float(dt),
           float(eplsion_k), D, H, W, s)
    out = [torch.empty_like(u) for _ in range(3)]
    L.call("hfm_pg_flux_f32", L.ptr_array([a.data_ptr() for a in vt]),
           L.ptr_array([a.data_ptr() for a in ks]), L.ptr_array([a.data_ptr() for a in out]),
           K[0].ctypes.data_as(L.C.c_void_p), D, H, W, s)
    return tuple(out)


def conv_sum(xs, kernels, scale: float = 1.0, source_bc: bool = False):
    """``scale * sum_q K_q (*) x_q`` (1-3 fields) in one pass; ``source_bc`` copies the
    x = 1 plane into x = 0 (boundary_condition_source)."""
    xs = [_prep(a) for a in xs]
    (D, H, W), _ = _spatial(xs[0])
    for t in xs:
        if t.shape != xs[0].shape:
            raise ValueError("all fields must share one shape")
    y = torch.empty_like(xs[0])
    K = np.ascontiguousarray(np.stack([kernel27(k) for k in kernels]), dtype=np.float32)
    L.call("hfm_conv27_sum_f32", L.ptr_array([a.data_ptr() for a in xs]), y.data_ptr(), len(xs),
           K.ctypes.data_as(L.C.c_void_p), float(scale), D, H, W,
           L.SUM_SOURCE_BC if source_bc else 0, L.stream_handle(xs[0].device))
    return y
