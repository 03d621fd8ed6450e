"""Multigrid-preconditioned FP64 conjugate-gradient solves of the anisotropic
resistivity forward problem ``div(sigma grad phi) = b``, batched.

The operator has the form of the reference's variable-coefficient layer
``A_operator_n`` (ART/BGS_multigrid_v1.ipynb#c40: a sum over the six
directions of a coefficient times ``x - x_nb`` with zero ghosts), stored as
three face-coefficient arrays per level.  The reference's V-cycle
(ART/Anisotropic_3d_Multigrid_7point.ipynb:402-454: mean restriction,
nearest prolongation, Jacobi smoothing) becomes a V(nu1, nu2) cycle with
weighted Jacobi, Galerkin coarse operators and a dense FP64 coarsest solve;
it preconditions an FP64 conjugate-gradient outer loop that owns the
residual ``r = b - A x`` and the solution update.  Inner storage is FP32 or
BF16 (FP32 arithmetic).

Structure:
  Pool      solver-state slots (FP64 scalars) and dot-product partial slots
            shared by every instance of one device, plus the parameters.
  Instance  one solve's resident hierarchy: faces, vectors and FP64 state.
  Plan      the work tables for a set of instances: one libhfm.so launch
            per (phase, level) covers all of them, so uneven sizes share
            launches and an instance's arithmetic is identical alone or in
            any mix (fixed-order per-instance reductions).
  MGBatch   a fixed batch (pool + instances + one plan), HIP-graph captured.
The serving loop (hfm.server) admits and retires instances between plans.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L

BLOCK = 256           # threads per block of the FCG kernels and the default elsewhere
CHUNK = 8 * BLOCK     # cells per block of the cell kernels (a thread owns 8)
RCHUNK = 4 * BLOCK    # coarse cells per block of the residual + restriction
BLOCKS = (128, 256, 512)   # threads per block a sweep / restriction launch may use
SCHUNK = 256          # cells per block of the batched setup kernels
PAIR_BLOCK = 512      # threads per block of the fused pass pairs (one tile each)
ALIGN = 8             # element alignment of sub-array offsets (16-byte rows)
ELEM = {"f32": 4, "bf16": 2, "f16": 2}
TORCH_STORE = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}
# outer (z, p) storage types built for each V-cycle storage type
OUTER_OF = {"f32": ("f32", "bf16"), "bf16": ("bf16", "f32"), "f16": ("f16", "f32")}


ILV_SEG = 512        # cells per lane-interleaved segment of the FP64 vectors


def ilv_len(n: int) -> int:
    """Length of an FP64 solver vector (x64, r64) of ``n`` cells: padded to a
    whole number of interleave segments."""
    return (n + ILV_SEG - 1) // ILV_SEG * ILV_SEG


def deinterleave(v: torch.Tensor, n: int, width: int) -> torch.Tensor:
    """Natural-order copy of an FP64 solver vector.

    For a grid whose row width is a multiple of 8 (the kernels' vector path)
    the vector is stored lane-interleaved: in each 512-cell segment, cells
    8 l + 2 q and 8 l + 2 q + 1 sit at offsets 128 q + 2 l and 128 q + 2 l + 1
    (a wave's four 16-byte accesses then cover contiguous bytes).  Other
    widths keep the natural order."""
    if width % 8:
        return v[:n].clone()
    return v.view(-1, 4, 64, 2).transpose(1, 2).reshape(-1)[:n]


def interleave(x: torch.Tensor, width: int) -> torch.Tensor:
    """Inverse of ``deinterleave``: the stored (padded) form of natural-order ``x``."""
    n = x.numel()
    out = torch.zeros(ilv_len(n), dtype=x.dtype, device=x.device)
    out[:n] = x.reshape(-1)
    if width % 8:
        return out
    return out.view(-1, 64, 4, 2).transpose(1, 2).reshape(-1)


def level_shapes(shape: Sequence[int], max_coarse_edge: int = 4) -> List[Tuple[int, int, int]]:
    """Halve every axis (rounding up) while any edge exceeds ``max_coarse_edge``."""
    shapes = [tuple(int(s) for s in shape)]
    while max(shapes[-1]) > max_coarse_edge:
        shapes.append(tuple((n + 1) // 2 for n in shapes[-1]))
    return shapes


def _round8(n: int) -> int:
    return (n + 7) // 8 * 8


def _align(n: int) -> int:
    return (n + ALIGN - 1) // ALIGN * ALIGN


def check_shape(shape, max_coarse_edge: int = 4):
    shape = tuple(int(v) for v in shape)
    if len(shape) != 3 or min(shape) < 1:
        raise ValueError(f"bad shape {shape}")
    lv = level_shapes(shape, max_coarse_edge)
    if len(lv) < 2:
        raise ValueError(f"grid {shape} is too small for a multigrid hierarchy")
    if math.prod(shape) >= 2 ** 31:
        raise ValueError(f"grid {shape} exceeds 2^31 cells")
    if math.prod(lv[-1]) > 64:
        raise ValueError(f"coarsest level {lv[-1]} exceeds 64 cells")
    return shape, lv


def fine_blocks(shape) -> int:
    return _round8((math.prod(shape) + CHUNK - 1) // CHUNK)


@dataclasses.dataclass
class Params:
    storage: str = "bf16"
    tol: float = 1e-8
    maxit: int = 200
    nu1: int = 2
    nu2: int = 2
    omega: float = 2.0 / 3.0
    alpha: float = 1.6
    max_coarse_edge: int = 4
    outer: Optional[str] = None   # z and p storage (default: ``storage``)
    # threads per block of the launches without reductions, keyed
    # ("sweep", mode, level) / ("restrict", level); default BLOCK (hfm.tune)
    blocks: Optional[dict] = None
    # levels (below the finest) of at most this many cells run in one
    # workgroup per instance (hfm_mg_coarse_cycle); the coarsest always does
    coarse_cells: int = 4096
    # fused pass pairs on levels that tile (hfm_mg_smooth_restrict: the last
    # pre-sweep with the restriction; hfm_mg_smooth2: the first two
    # post-sweeps); results are bitwise those of the separate launches.  Only
    # the finest `fuse_levels` levels, each where the batch has enough tiles
    # there: a fused pair walks its tiles plane by plane, which pays on the
    # two finest levels of a large batch (MI355X, 512-instance batch: L0 1.11
    # + 1.10 ms fused vs 1.66 + 1.67 separate; L1 0.17 + 0.17 vs 0.23 + 0.21)
    fuse: bool = True
    fuse_levels: int = 2
    # an MGBatch fuses the fine level only with at least this many tiles over
    # its instances (a few tiles leave most CUs idle: one 64 x 256 x 256
    # instance has 86).  The choice is the batch's (its pool's), not each
    # plan's: every plan of the batch (the subsets as instances finish, the
    # stream groups) then runs the same kernels and gives bitwise the same
    # results.  (The fused post pair sums the r.z partials per tile, the
    # separate sweep per block of cells: an instance in a batch that fuses
    # and alone in one that does not agrees to rounding in that sum.)
    fuse_min_tiles: int = 256
    # outer iteration: "fcg" (flexible CG preconditioned by one V-cycle) or
    # "refine" (FP64 defect correction x += V(r): the stationary V-cycle
    # iteration with an FP64 residual and solution, ART/Anisotropic_3d_Multigrid_
    # 7point.ipynb:402-454's cycle loop stopped at a tolerance)
    method: str = "fcg"
    # planes per tile of the fused pass pairs (at most; hfm_mg_tile): fewer
    # planes give more tiles (blocks) at the cost of more halo planes
    tile_planes: int = 32
    # FP16 storage: every iteration scales the fine residual the V-cycle gets
    # by a power of 2 so that its RMS is near 2^f16_rms_exp (the V-cycle's
    # values then stay inside FP16's range; FCG is invariant to the scale)
    f16_rms_exp: int = -4

    def __post_init__(self):
        if self.storage not in ELEM:
            raise ValueError(f"storage must be one of {sorted(ELEM)}")
        if self.outer is None:
            self.outer = self.storage
        if self.outer not in OUTER_OF[self.storage]:
            raise ValueError(f"outer storage for {self.storage} must be one of "
                             f"{OUTER_OF[self.storage]}")
        if self.nu1 < 1 or self.nu2 < 1:
            raise ValueError("nu1 and nu2 must be >= 1")
        if self.method not in L.METHOD:
            raise ValueError(f"method must be one of {sorted(L.METHOD)}")
        if self.tile_planes < 2:
            raise ValueError("tile_planes must be >= 2")


class _Arena:
    """First-fit allocator of integer ranges (dot-partial slots)."""

    def __init__(self, size: int):
        self.size = size
        self.free = [(0, size)]

    def alloc(self, n: int) -> int:
        for q, (off, sz) in enumerate(self.free):
            if sz >= n:
                if sz == n:
                    self.free.pop(q)
                else:
                    self.free[q] = (off + n, sz - n)
                return off
        raise MemoryError("dot-slot arena exhausted")

    def release(self, off: int, n: int):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, s in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + s)
            else:
                merged.append((o, s))
        self.free = merged


class Pool:
    """Device-wide solver state: per-slot FP64 scalars and dot-product partials."""

    def __init__(self, capacity: int, dot_slots: int, params: Params, device=None, alloc=None):
        """``alloc(numel, dtype)`` returns a zeroed device tensor for every
        buffer of the pool and its instances (default ``torch.zeros``); tests
        pass one that surrounds each buffer with guard bands."""
        if capacity < 1 or capacity > 65535:
            raise ValueError("pool capacity must be in [1, 65535]")
        self.params = params
        self.device = torch.device(device or "cuda")
        self.zeros = alloc if alloc is not None else \
            (lambda n, dtype: torch.zeros(n, dtype=dtype, device=self.device))
        self.plan_only = self.device.type != "cuda"
        if not self.plan_only:
            if not torch.cuda.is_available():
                raise RuntimeError("the solver needs a GPU (the HIP library has no CPU path)")
            L.load()
        self.store_id = L.STORE[params.storage]
        self.outer_id = L.STORE[params.outer]
        # fine-level pass pairs fused for every plan of this pool (the batch or
        # server that owns the pool may turn it off: see Params.fuse_min_tiles),
        # and those of the levels below it (up to Params.fuse_levels)
        self.fuse_fine = params.fuse
        self.fuse_coarse = params.fuse
        self.capacity = capacity
        self.sc = self.zeros(capacity * L.SCALARS_DTYPE.itemsize, torch.uint8)
        self.dots = self.zeros(max(dot_slots, 1), torch.float64)
        self._free_slots = list(range(capacity - 1, -1, -1))
        self._dots = _Arena(max(dot_slots, 1))

    def acquire(self, nblk: int) -> Tuple[int, int]:
        if not self._free_slots:
            raise MemoryError("no free solver slot")
        off = self._dots.alloc(nblk)
        return self._free_slots.pop(), off

    def release(self, slot: int, dots_off: int, nblk: int):
        self._free_slots.append(slot)
        self._dots.release(dots_off, nblk)

    def free_slots(self) -> int:
        return len(self._free_slots)

    def scalars(self) -> np.ndarray:
        return self.sc.cpu().numpy().view(L.SCALARS_DTYPE).copy()

    def active_view(self):
        return self.sc.view(torch.int32).view(self.capacity, -1)[:, 18]

    def xlag_view(self):
        return self.sc.view(torch.int32).view(self.capacity, -1)[:, 22]

    def stage(self, insts: Sequence["Instance"]):
        """Mark the instances' slots for (re)initialisation (active = 1, iter = -1)."""
        if not insts:
            return
        p = self.params
        rec = np.zeros(len(insts), dtype=L.SCALARS_DTYPE)
        rec["tol"], rec["maxit"], rec["active"], rec["iter"] = p.tol, p.maxit, 1, -1
        rec["method"] = L.METHOD[p.method]
        rec["slot0"] = [i.dots_off for i in insts]
        rec["nslot"] = [i.nblk for i in insts]
        if p.storage == "f16":
            # the first residual (b) scaled to an RMS near 2^f16_rms_exp; the
            # finalize keeps it there as the residual falls
            norms = torch.stack([torch.linalg.vector_norm(i.b32.double()) for i in insts])
            norms = norms.cpu().numpy()
            sbase = [p.f16_rms_exp + int(round(0.5 * math.log2(max(i.N, 1)))) for i in insts]
            rec["sbase"], rec["scaled"] = sbase, 1
            rec["sexp"] = [sb - (math.frexp(float(bn))[1] - 1)
                           if bn > 0 and math.isfinite(bn) else sb
                           for sb, bn in zip(sbase, norms)]
        rows = torch.from_numpy(rec.view(np.uint8).reshape(len(insts), -1).copy()).to(self.device)
        idx = torch.tensor([i.slot for i in insts], device=self.device)
        self.sc.view(self.capacity, -1).index_copy_(0, idx, rows)

    def finalize(self, kind: str, s, slots: Optional[Tuple[int, int]] = None):
        """Finalize step ``kind`` for the slots [lo, hi) (default: all): a plan
        whose launches run beside another plan's on a second stream touches
        only its own slots."""
        lo, hi = slots if slots is not None else (0, self.capacity)
        L.call("hfm_fcg_finalize", self.sc.data_ptr() + lo * L.SCALARS_DTYPE.itemsize, hi - lo,
               self.dots.data_ptr(), L.FIN[kind], s)


def iteration_bytes(inst, params: Params) -> float:
    """Algorithmic bytes of one FCG iteration of ``inst``: on every level but the
    coarsest, nu1 + nu2 sweeps (x, b, 3 faces read, x written) and the residual
    restriction (x, b, 3 faces read, the coarse rhs written); on the fine level
    the direction pass (z, p read, 3 FP32 faces, p written) and the update (p,
    3 FP32 faces, r64 read and written, rlow and the first sweep written; x64
    and the previous direction on every other iteration: their mean)."""
    e, eo = ELEM[params.storage], ELEM[params.outer]
    nu = params.nu1 + params.nu2
    tot = sum(D * H * W * (nu * 6 * e + 5 * e + e / 8) for (D, H, W) in inst.levels[:-1])
    upd = 2 * e + eo + 12 + 16 + 0.5 * (16 + eo)
    return tot + math.prod(inst.levels[0]) * (3 * eo + 12 + upd)


class Instance:
    """One solve's resident hierarchy and FP64 solver state."""

    def __init__(self, pool: Pool, shape):
        self.pool = pool
        p = pool.params
        self.shape, self.levels = check_shape(shape, p.max_coarse_edge)
        st = TORCH_STORE[p.storage]
        zeros = pool.zeros
        self.face_off, self.face32, self.faceT = [], [], []
        self.X, self.Y, self.B = [], [], []
        for (D, H, W) in self.levels:
            N = D * H * W
            offs, tot = [], 0
            for sz in (N, N, N, D * H, D * W, H * W):
                offs.append(tot)
                tot += _align(sz)
            self.face_off.append(offs)
            f32 = zeros(tot, torch.float32)
            self.face32.append(f32)
            self.faceT.append(f32 if p.storage == "f32" else zeros(tot, st))
            self.X.append(zeros(N, st))
            self.Y.append(zeros(N, st))
            self.B.append(zeros(N, st))
        N0 = math.prod(self.shape)
        self.N = N0
        self.sigma = zeros(3 * _align(N0), torch.float32)
        self.b32 = zeros(N0, torch.float32)
        self.x64 = zeros(ilv_len(N0), torch.float64)
        self.r64 = zeros(ilv_len(N0), torch.float64)
        so = TORCH_STORE[p.outer]
        self.P = [zeros(N0, so), zeros(N0, so)]
        # z = V(rlow) has a buffer of its own (type outer): the fine level's last
        # post-sweep writes it whether or not its pass pair runs fused
        self.Z = zeros(N0, so)
        nc = math.prod(self.levels[-1])
        self.ainv = zeros(nc * nc, torch.float64)
        self.nblk = fine_blocks(self.shape)
        self.slot, self.dots_off = pool.acquire(self.nblk)
        self.released = False

    def release(self):
        """Return the slot and let the caching allocator reuse the buffers."""
        if not self.released:
            self.pool.release(self.slot, self.dots_off, self.nblk)
            self.released = True
            for name in ("face32", "faceT", "X", "Y", "B", "P"):
                setattr(self, name, [])
            self.sigma = self.b32 = self.r64 = self.ainv = self.Z = None
            self._tpl = None      # its addresses are no longer valid

    def nbytes(self) -> int:
        ts = [*self.face32, *self.X, *self.Y, *self.B, *self.P, self.sigma, self.b32, self.x64,
              self.r64, self.ainv, self.Z]
        if self.pool.params.storage != "f32":
            ts += self.faceT
        return sum(t.numel() * t.element_size() for t in ts)

    @staticmethod
    def estimate_bytes(shape, params: Params) -> int:
        """Device bytes an instance of ``shape`` holds (admission accounting)."""
        lv = level_shapes(shape, params.max_coarse_edge)
        e = ELEM[params.storage]
        tot = 0
        for (D, H, W) in lv:
            N = D * H * W
            faces = sum(_align(s) for s in (N, N, N, D * H, D * W, H * W))
            tot += faces * 4 + (faces * e if params.storage != "f32" else 0) + 3 * N * e
        N0 = math.prod(shape)
        nc = math.prod(lv[-1])
        eo = ELEM[params.outer]
        return tot + 3 * _align(N0) * 4 + N0 * (4 + 3 * eo) + 16 * ilv_len(N0) + nc * nc * 8

    def faces(self, lev: int, low: bool) -> List[int]:
        """Addresses of the six face arrays of a level (storage copy if ``low``)."""
        buf = self.faceT[lev] if low else self.face32[lev]
        es = buf.element_size()
        return [buf.data_ptr() + o * es for o in self.face_off[lev]]

    def stage_sigma(self, sx, sy, sz, b):
        """Keep the cell conductivities resident and load the right-hand side."""
        N, N8 = self.N, _align(self.N)
        for q, a in enumerate((sx, sy, sz)):
            a = torch.as_tensor(a).to(device=self.pool.device, dtype=torch.float32).reshape(-1)
            if a.numel() != N:
                raise ValueError("conductivity size mismatch")
            self.sigma[q * N8:q * N8 + N].copy_(a)
        b = torch.as_tensor(b).to(device=self.pool.device, dtype=torch.float32).reshape(-1)
        if b.numel() != N:
            raise ValueError("rhs size mismatch")
        self.b32.copy_(b)

    def set_faces(self, faces, b):
        """Load FP32 face arrays (cx, cy, cz, bxh, byh, bzh) and the rhs directly."""
        D, H, W = self.shape
        shapes = [(D, H, W)] * 3 + [(D, H), (D, W), (H, W)]
        for q, (a, shp) in enumerate(zip(faces, shapes)):
            a = torch.as_tensor(a).to(device=self.pool.device, dtype=torch.float32).reshape(-1)
            if a.numel() != math.prod(shp):
                raise ValueError(f"face array {q} has {a.numel()} values, expected {shp}")
            o = self.face_off[0][q]
            self.face32[0][o:o + a.numel()].copy_(a)
        b = torch.as_tensor(b).to(device=self.pool.device, dtype=torch.float32).reshape(-1)
        if b.numel() != self.N:
            raise ValueError("rhs size mismatch")
        self.b32.copy_(b)

    def solution(self) -> torch.Tensor:
        """FP64 solution, (D,H,W), natural order (a copy of the solver vector)."""
        return deinterleave(self.x64, self.N, self.shape[2]).view(self.shape)


@dataclasses.dataclass
class _Layout:
    """Block layout of one launch geometry."""
    blk0: np.ndarray
    nblk: np.ndarray
    blk_task: torch.Tensor  # uint16 per block (device)
    nblocks: int


def _layout(device, sizes, chunk) -> _Layout:
    sizes = np.asarray(sizes, dtype=np.int64)
    nblk = ((sizes + chunk - 1) // chunk + 7) // 8 * 8
    blk0 = np.concatenate([[0], np.cumsum(nblk)[:-1]]).astype(np.int64)
    bt = np.repeat(np.arange(len(sizes), dtype=np.uint16), nblk)
    return _Layout(blk0, nblk, torch.from_numpy(bt.view(np.int16).copy()).to(device),
                   int(nblk.sum()))


def _record(inst: Instance, lev: int, coarse=None):
    t = np.zeros((), dtype=L.TASK_DTYPE)
    D, H, W = inst.levels[lev]
    t["D"], t["H"], t["W"], t["inst"] = D, H, W, inst.slot
    t["slot0"] = inst.dots_off
    t["magW"], t["magH"] = L.magic(W), L.magic(H)
    if coarse is not None:
        Dc, Hc, Wc = coarse
        t["Dc"], t["Hc"], t["Wc"] = Dc, Hc, Wc
        t["magWc"], t["magHc"] = L.magic(Wc), L.magic(Hc)
    return t


def chain_start(levels, coarse_cells: int) -> int:
    """First level of an instance's one-launch coarse cycle (at least 1)."""
    for l in range(1, len(levels)):
        if math.prod(levels[l]) <= coarse_cells:
            return l
    return len(levels) - 1


def chain_records(inst: "Instance", lc: int) -> list:
    """Records of levels lc .. coarsest for hfm_mg_coarse_cycle."""
    recs = []
    L_ = len(inst.levels)
    for l in range(lc, L_):
        t = _record(inst, l, coarse=inst.levels[l + 1] if l + 1 < L_ else None)
        t["c"] = inst.faces(l, True)
        t["xin"], t["xout"], t["b"] = inst.X[l].data_ptr(), inst.Y[l].data_ptr(), inst.B[l].data_ptr()
        t["ec"] = inst.B[l + 1].data_ptr() if l + 1 < L_ else inst.ainv.data_ptr()
        recs.append(t)
    return recs


def _tile_table(device, recs, ntiles):
    """Task table + tile layout of a fused pass-pair launch: task q owns tiles
    blk0 .. blk0 + ntiles[q] - 1, numbered contiguously; the tile -> task map
    is padded with 0xFFFF to a multiple of 8 (hfm_mg_smooth_restrict)."""
    arr = np.array(recs, dtype=L.TASK_DTYPE)
    n = np.asarray(ntiles, dtype=np.int64)
    blk0 = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    arr["blk0"], arr["nblk"] = blk0, n
    total = int(n.sum())
    padded = _round8(total)
    bt = np.full(padded, 0xFFFF, dtype=np.uint16)
    bt[:total] = np.repeat(np.arange(len(n), dtype=np.uint16), n)
    lay = _Layout(blk0, n, torch.from_numpy(bt.view(np.int16).copy()).to(device), padded)
    return torch.from_numpy(arr.view(np.uint8).copy()).to(device), lay


def _table(device, recs, sizes, chunk):
    """Device task table + block layout for the records (one per member)."""
    arr = np.array(recs, dtype=L.TASK_DTYPE)
    lay = _layout(device, sizes, chunk)
    arr["blk0"], arr["nblk"] = lay.blk0, lay.nblk
    return torch.from_numpy(arr.view(np.uint8).copy()).to(device), lay


def fused_levels(inst: Instance, p: Params, lc: int):
    """Levels (below ``lc``, where the coarse cycle starts) whose pass pairs run
    fused: ({lev: tile} for smooth + restrict, {lev: tile} for the two
    post-sweeps).  The fine level's post pair carries the r.z partials, one
    per tile, so its tiles must fit the instance's dot slots."""
    sr, s2 = {}, {}
    if not p.fuse:
        return sr, s2
    for lev in range(min(lc, len(inst.levels) - 1, p.fuse_levels)):
        if not (inst.pool.fuse_fine if lev == 0 else inst.pool.fuse_coarse):
            continue
        tl = L.tile(inst.levels[lev], p.tile_planes)
        if tl is None:
            continue
        if p.nu1 >= 2:
            sr[lev] = tl
        if p.nu2 >= 2 and not (lev == 0 and p.nu2 == 2 and tl[2] > inst.nblk):
            s2[lev] = tl
    return sr, s2


def instance_templates(inst: Instance, p: Params, lc: int) -> dict:
    """Task records of every launch of the FCG iteration that touch ``inst``,
    keyed by launch; block fields are filled in when a plan is assembled.
    ``lc``: the instance's first coarse-cycle level (levels below it have
    per-level launches, and may have fused pass pairs)."""
    sr, s2 = fused_levels(inst, p, lc)

    def xy(k, lev):  # sweeps alternate between the X and Y buffers
        return (inst.X, inst.Y)[k % 2][lev].data_ptr()

    def post_out(lev, s):
        # buffer written by post-sweep s; a fused pair (s = 0, 1) writes the
        # buffer of s = 0, since its input (pre-sweep nu1 - 1) is the other one
        return xy(p.nu1 + s - (1 if lev in s2 and s >= 1 else 0), lev)

    def E(lev):  # buffer holding a level's correction after its cycle
        if lev == len(inst.levels) - 1:
            return inst.X[lev].data_ptr()
        return post_out(lev, p.nu2 - 1)

    tpl = {}
    for lev in range(len(inst.levels) - 1):
        for s in range(p.nu1):
            t = _record(inst, lev)
            t["c"] = inst.faces(lev, True)
            t["xin"] = xy(s - 1, lev) if s else 0
            t["xout"] = xy(s, lev)
            t["b"] = inst.B[lev].data_ptr()
            tpl[("down", lev, s)] = t
        t = _record(inst, lev, coarse=inst.levels[lev + 1])
        t["c"] = inst.faces(lev, True)
        t["xin"] = xy(p.nu1 - 1, lev)
        t["b"] = inst.B[lev].data_ptr()
        t["xout"] = inst.B[lev + 1].data_ptr()
        tpl[("restrict", lev)] = t
        for s in range(p.nu2):
            t = _record(inst, lev, coarse=inst.levels[lev + 1])
            t["c"] = inst.faces(lev, True)
            t["xin"] = xy(p.nu1 - 1, lev) if s == 0 else post_out(lev, s - 1)
            t["xout"] = post_out(lev, s)
            t["b"] = inst.B[lev].data_ptr()
            if s == 0:
                t["ec"] = E(lev + 1)
            if lev == 0 and s == p.nu2 - 1:
                t["xout"] = inst.Z.data_ptr()
            tpl[("up", lev, s)] = t
        if lev in sr:
            t = tpl[("restrict", lev)].copy()
            last = tpl[("down", lev, p.nu1 - 1)]
            t["xin"], t["xout"] = last["xin"], last["xout"]
            t["rlow"] = inst.B[lev + 1].data_ptr()
            t["tile_rows"], t["tile_planes"] = sr[lev][:2]
            tpl[("smooth_restrict", lev)] = t
        if lev in s2:
            t = tpl[("up", lev, 0)].copy()
            t["xout"] = tpl[("up", lev, 1)]["xout"]
            t["tile_rows"], t["tile_planes"] = s2[lev][:2]
            tpl[("smooth2", lev)] = t
    inst._tiles = {lev: tl[2] for lev, tl in {**sr, **s2}.items()}
    lc = len(inst.levels) - 1
    t = _record(inst, lc)
    t["b"] = inst.B[lc].data_ptr()
    t["xout"] = inst.X[lc].data_ptr()
    t["ec"] = inst.ainv.data_ptr()
    tpl[("coarse",)] = t
    base = _record(inst, 0)
    base["c"] = inst.faces(0, False)
    base["x64"], base["r64"] = inst.x64.data_ptr(), inst.r64.data_ptr()
    base["rlow"] = inst.B[0].data_ptr()
    base["b32"] = inst.b32.data_ptr()
    base["b"] = tpl[("up", 0, p.nu2 - 1)]["xout"]   # z: the fine level's last post-sweep
    base["xout"] = xy(0, 0)     # first fine-level pre-sweep, fused into init / update
    tpl[("init",)] = base
    for par in range(2):
        d = base.copy()
        d["xin"], d["xout"] = inst.P[par].data_ptr(), inst.P[1 - par].data_ptr()
        tpl[("dir", par)] = d
        u = base.copy()
        u["xin"] = inst.P[1 - par].data_ptr()
        u["ec"] = inst.P[par].data_ptr()      # the previous direction (deferred x update)
        tpl[("update", par)] = u
    return tpl


class SetupPlan:
    """Work tables that build the hierarchies of a set of instances."""

    def __init__(self, pool: Pool, insts: Sequence[Instance]):
        self.pool = pool
        dev = pool.device
        self.n = len(insts)
        recs = []
        for inst in insts:
            t = _record(inst, 0)
            N8 = _align(inst.N) * 4
            base = inst.sigma.data_ptr()
            t["xin"], t["b"], t["ec"] = base, base + N8, base + 2 * N8
            t["o"] = inst.faces(0, False)
            recs.append(t)
        self.faces = _table(dev, recs, [i.N for i in insts], CHUNK)
        self.galerkin = []
        nlev = max(len(i.levels) for i in insts)
        for lev in range(nlev - 1):
            members = [i for i in insts if len(i.levels) > lev + 1]
            recs = []
            for inst in members:
                t = _record(inst, lev)
                t["c"] = inst.faces(lev, False)
                t["o"] = inst.faces(lev + 1, False)
                recs.append(t)
            self.galerkin.append(_table(dev, recs, [math.prod(i.levels[lev + 1]) for i in members],
                                        SCHUNK))
        self.convert = None
        if pool.params.storage != "f32":
            recs, sizes = [], []
            for inst in insts:
                for lev in range(len(inst.levels)):
                    t = np.zeros((), dtype=L.TASK_DTYPE)
                    t["D"], t["H"], t["W"] = inst.face32[lev].numel(), 1, 1
                    t["xin"], t["xout"] = inst.face32[lev].data_ptr(), inst.faceT[lev].data_ptr()
                    recs.append(t)
                    sizes.append(inst.face32[lev].numel())
            self.convert = _table(dev, recs, sizes, CHUNK)   # 8 elements per thread
        recs = []
        for inst in insts:
            lc = len(inst.levels) - 1
            t = _record(inst, lc)
            t["c"] = inst.faces(lc, False)
            t["xout"] = inst.ainv.data_ptr()
            recs.append(t)
        self.inverse = torch.from_numpy(np.array(recs, dtype=L.TASK_DTYPE).view(np.uint8)
                                        .copy()).to(dev)

    def build_faces(self, s, inv_h2: float = 1.0):
        tasks, lay = self.faces
        L.call("hfm_mg_faces_batched", tasks.data_ptr(), lay.blk_task.data_ptr(), lay.nblocks,
               float(inv_h2), s)

    def build_hierarchy(self, s):
        for tasks, lay in self.galerkin:
            L.call("hfm_mg_galerkin_batched", tasks.data_ptr(), lay.blk_task.data_ptr(),
                   lay.nblocks, s)
        if self.convert is not None:
            tasks, lay = self.convert
            L.call("hfm_mg_convert_batched", tasks.data_ptr(), lay.blk_task.data_ptr(),
                   lay.nblocks, self.pool.store_id, s)
        L.call("hfm_mg_coarse_inverse", self.inverse.data_ptr(), self.n, s)


class Plan:
    """Launch tables of the FCG iteration for a set of instances.

    Assembled from each instance's cached task records, so rebuilding a plan
    when the set changes costs a few numpy concatenations and uploads.
    """

    def __init__(self, pool: Pool, insts: Sequence[Instance],
                 slots: Optional[Tuple[int, int]] = None):
        if not insts:
            raise ValueError("empty plan")
        self.pool = pool
        self.insts = list(insts)
        # finalize range: every slot of the plan's instances (default: the pool)
        self.slots = slots
        if slots is not None and not all(slots[0] <= i.slot < slots[1] for i in insts):
            raise ValueError("plan instances outside its slot range")
        p = pool.params
        dev = pool.device
        lcs = {i.slot: chain_start(i.levels, p.coarse_cells) for i in insts}
        for inst in insts:
            # records depend on where the coarse cycle starts (fused levels
            # end there) and on fusion itself
            key = (lcs[inst.slot], p.fuse, p.fuse_levels, pool.fuse_fine, pool.fuse_coarse,
                   p.nu1, p.nu2, p.tile_planes)
            if getattr(inst, "_tpl", None) is None or inst._tpl_key != key:
                inst._tpl = instance_templates(inst, p, lcs[inst.slot])
                inst._tpl_key = key
        nlev = max(len(i.levels) for i in insts)
        self.nlev = nlev

        def op(key, chunk, size_of, lev=None, keep=None):
            # per-level launches stop where an instance's coarse cycle starts
            members = [i for i in insts if key in i._tpl and (lev is None or lev < lcs[i.slot])
                       and (keep is None or keep(i))]
            if not members:
                return None
            recs, sizes = [i._tpl[key] for i in members], [size_of(i) for i in members]
            if chunk is None:   # fused pass pair: one block per tile
                return _tile_table(dev, recs, sizes)
            return _table(dev, recs, sizes, chunk)

        def fused(key):
            """Members whose pass pair ``key`` runs fused (fused_levels)."""
            return lambda i: key in i._tpl

        blocks = p.blocks or {}
        self.ops_down, self.ops_up = [], []
        for lev in range(nlev - 1):
            cells = lambda i, lev=lev: math.prod(i.levels[lev])  # noqa: E731
            tiles = lambda i, lev=lev: i._tiles[lev]  # noqa: E731
            fr, f2 = fused(("smooth_restrict", lev)), fused(("smooth2", lev))
            not_fr = lambda i, fr=fr: not fr(i)  # noqa: E731
            not_f2 = lambda i, f2=f2: not f2(i)  # noqa: E731
            for s in range(p.nu1):
                if lev == 0 and s == 0:
                    continue    # computed by the FCG update (and init) kernels
                mode = "zero" if s == 0 else "plain"
                b = blocks.get(("sweep", mode, lev), BLOCK)
                tab = op(("down", lev, s), 8 * b, cells, lev, not_fr if s == p.nu1 - 1 else None)
                if tab:
                    self.ops_down.append(("sweep", L.SWEEP[mode], tab[0], tab[1], False, lev, b))
            b = blocks.get(("restrict", lev), BLOCK)
            tab = op(("restrict", lev), 4 * b, lambda i, lev=lev: math.prod(i.levels[lev + 1]),
                     lev, not_fr)
            if tab:
                self.ops_down.append(("restrict", 0, tab[0], tab[1], False, lev, b))
            tab = op(("smooth_restrict", lev), None, tiles, lev, fr)
            if tab:
                self.ops_down.append(("smooth_restrict", 0, tab[0], tab[1], False, lev, PAIR_BLOCK))
            ups = []
            tab = op(("smooth2", lev), None, tiles, lev, f2)
            if tab:
                ups.append(("smooth2", 0, tab[0], tab[1], lev == 0 and p.nu2 == 2, lev,
                            PAIR_BLOCK))
            for s in range(p.nu2):
                mode = "correct" if s == 0 else "plain"
                dot = lev == 0 and s == p.nu2 - 1
                b = BLOCK if dot else blocks.get(("sweep", mode, lev), BLOCK)
                tab = op(("up", lev, s), 8 * b, cells, lev, not_f2 if s < 2 else None)
                if tab:
                    ups.append(("sweep", L.SWEEP[mode], tab[0], tab[1], dot, lev, b))
            self.ops_up.insert(0, ups)
        self.coarse_tasks = torch.from_numpy(
            np.array([i._tpl[("coarse",)] for i in insts], dtype=L.TASK_DTYPE).view(np.uint8)
            .copy()).to(dev)
        recs, chain = [], []
        for i in insts:
            r = chain_records(i, lcs[i.slot])
            chain += [len(recs), len(r)]
            recs += r
        self.chain_recs = torch.from_numpy(np.array(recs, dtype=L.TASK_DTYPE).view(np.uint8)
                                           .copy()).to(dev)
        self.chain = torch.tensor(chain, dtype=torch.int32).to(dev)
        self.chain_levels = sorted(set(lcs.values()))
        fine = lambda i: i.N  # noqa: E731
        self.init_tasks, self.flay = op(("init",), CHUNK, fine)
        self.dir_tasks = [op(("dir", par), CHUNK, fine)[0] for par in range(2)]
        self.upd_tasks = [op(("update", par), CHUNK, fine)[0] for par in range(2)]
        self._graphs = {}

    def _launch(self, op, s):
        kind, mode, tasks, lay, dot, _, block = op
        pool, p = self.pool, self.pool.params
        if kind == "sweep":
            # the sweep carrying the r.z dot is the fine level's last: it writes z
            out = pool.outer_id if dot else pool.store_id
            L.call("hfm_mg_sweep", tasks.data_ptr(), lay.blk_task.data_ptr(), lay.nblocks,
                   pool.sc.data_ptr(), pool.store_id, out, mode, p.omega, p.alpha, block,
                   pool.dots.data_ptr() if dot else 0, s)
        elif kind == "smooth2":
            out = pool.outer_id if dot else pool.store_id
            L.call("hfm_mg_smooth2", tasks.data_ptr(), lay.blk_task.data_ptr(), lay.nblocks,
                   pool.sc.data_ptr(), pool.store_id, out, p.omega, p.alpha,
                   pool.dots.data_ptr() if dot else 0, s)
        elif kind == "smooth_restrict":
            L.call("hfm_mg_smooth_restrict", tasks.data_ptr(), lay.blk_task.data_ptr(),
                   lay.nblocks, pool.sc.data_ptr(), pool.store_id, p.omega, s)
        else:
            L.call("hfm_mg_resid_restrict", tasks.data_ptr(), lay.blk_task.data_ptr(),
                   lay.nblocks, pool.sc.data_ptr(), pool.store_id, block, s)

    def init(self, s):
        """x = 0, r = b for the plan's staged instances, then finalize INIT."""
        pool, fl = self.pool, self.flay
        L.call("hfm_fcg_init_sweep", self.init_tasks.data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, pool.sc.data_ptr(), pool.store_id, pool.params.omega,
               pool.dots.data_ptr(), s)
        pool.finalize("init", s, self.slots)

    def precondition(self, s):
        """z = V(rlow): one V-cycle over every instance of the plan."""
        for op in self.ops_down:
            self._launch(op, s)
        self.launch_coarse_cycle(s)
        for ups in self.ops_up:
            for op in ups:
                self._launch(op, s)

    def launch_coarse_cycle(self, s):
        """The V-cycle below each instance's chain start, one workgroup per instance."""
        p = self.pool.params
        L.call("hfm_mg_coarse_cycle", self.chain_recs.data_ptr(), self.chain.data_ptr(),
               len(self.insts), self.pool.sc.data_ptr(), self.pool.store_id, p.nu1, p.nu2,
               p.omega, p.alpha, s)

    def launch_dir(self, parity, s):
        pool, fl = self.pool, self.flay
        L.call("hfm_fcg_dir", self.dir_tasks[parity].data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, pool.sc.data_ptr(), pool.outer_id, pool.dots.data_ptr(), s)

    def launch_update(self, parity, s):
        pool, fl = self.pool, self.flay
        L.call("hfm_fcg_update_sweep", self.upd_tasks[parity].data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, pool.sc.data_ptr(), pool.store_id, pool.outer_id, pool.params.omega,
               pool.dots.data_ptr(), s)

    def launch_flush(self, parity, s):
        pool, fl = self.pool, self.flay
        L.call("hfm_fcg_flush", self.upd_tasks[parity].data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, pool.sc.data_ptr(), pool.outer_id, s)

    def iteration(self, s, parity):
        """One FCG iteration for every active instance of the plan; instances
        that stop get the pending half of their deferred solution update."""
        self.precondition(s)
        self.pool.finalize("beta", s, self.slots)
        self.launch_dir(parity, s)
        self.pool.finalize("alpha", s, self.slots)
        self.launch_update(parity, s)
        self.pool.finalize("check", s, self.slots)
        self.launch_flush(parity, s)

    def graph(self, m: int):
        """A HIP graph of ``m`` (rounded up to even) iterations, captured once per plan."""
        m += m % 2
        if m not in self._graphs:
            dev = self.pool.device
            side = torch.cuda.Stream(dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=side):
                s = L.stream_handle(dev)
                for it in range(m):
                    self.iteration(s, it % 2)
            torch.cuda.current_stream(dev).wait_stream(side)
            self._graphs[m] = g
        return self._graphs[m], m


class PersistentProgram:
    """``m`` FCG iterations of a plan as one hfm_mg_persistent launch.

    The phases are the plan's launches in the order ``Plan.iteration`` makes
    them (V-cycle down, coarse cycle, up, finalize BETA, direction, ALPHA,
    update, CHECK, flush), with the same tables; the cell kernels run with
    256-thread blocks (block size changes no value: see ``hfm.tune``).  The
    per-(phase, slot) item counts and each phase's predecessor per slot are
    computed here from the tables, once per program."""

    def __init__(self, plan: "Plan", m: int):
        pool = plan.pool
        p = pool.params
        self.pool, self.plan, self.m = pool, plan, int(m)
        dev = pool.device
        lo, hi = plan.slots if plan.slots is not None else (0, pool.capacity)
        nslot = pool.capacity
        dots = pool.dots.data_ptr()
        rows, item_slots = [], []

        def host_slots(tasks, lay, tiles=False):
            # the instance slot of each item (block) g; a fused pair's block g
            # runs tile (g % 8) (n / 8) + g / 8 (the kernels' XCD map)
            inst = tasks.cpu().numpy().view(L.TASK_DTYPE)["inst"].astype(np.int64)
            bt = lay.blk_task.cpu().numpy().view(np.uint16).astype(np.int64)
            if tiles:
                g = np.arange(len(bt))
                bt = bt[(g & 7) * (len(bt) >> 3) + (g >> 3)]
            return np.where(bt == 0xFFFF, -1, inst[np.minimum(bt, len(inst) - 1)])

        def add(kind, nitems, slots, tasks=0, blk=0, mode=0, dot=0, chain=0, slot0=0):
            r = np.zeros((), dtype=L.PHASE_DTYPE)
            r["kind"], r["mode"], r["dot"], r["nitems"] = L.PH[kind], mode, int(bool(dot)), nitems
            r["nu1"], r["nu2"], r["slot0"] = p.nu1, p.nu2, slot0
            r["omega"], r["alpha"] = p.omega, p.alpha
            r["tasks"], r["blk_task"], r["chain"], r["dots"] = tasks, blk, chain, dots
            rows.append(r)
            item_slots.append(slots)

        def add_op(op):
            kind, mode, tasks, lay, dot, _, _ = op
            tiles = kind in ("smooth_restrict", "smooth2")
            add(kind, lay.nblocks, host_slots(tasks, lay, tiles), tasks.data_ptr(),
                lay.blk_task.data_ptr(), mode=mode, dot=dot)

        fin_slots = np.arange(lo, hi, dtype=np.int64)
        chain = plan.chain.cpu().numpy().reshape(-1, 2)
        crec = plan.chain_recs.cpu().numpy().view(L.TASK_DTYPE)["inst"].astype(np.int64)
        coarse_slots = np.where(chain[:, 1] > 0, crec[np.minimum(chain[:, 0], len(crec) - 1)], -1)
        fine = [host_slots(plan.dir_tasks[q], plan.flay) for q in range(2)]
        for it in range(self.m):
            par = it % 2
            for op in plan.ops_down:
                add_op(op)
            add("coarse_cycle", len(plan.insts), coarse_slots, plan.chain_recs.data_ptr(),
                chain=plan.chain.data_ptr())
            for ups in plan.ops_up:
                for op in ups:
                    add_op(op)
            for kind, fin, tasks in (("finalize", "beta", None), ("dir", None, plan.dir_tasks[par]),
                                     ("finalize", "alpha", None),
                                     ("update", None, plan.upd_tasks[par]),
                                     ("finalize", "check", None),
                                     ("flush", None, plan.upd_tasks[par])):
                if kind == "finalize":
                    add("finalize", hi - lo, fin_slots, mode=L.FIN[fin], slot0=lo)
                else:
                    add(kind, plan.flay.nblocks, fine[par], tasks.data_ptr(),
                        plan.flay.blk_task.data_ptr())
        self.nphase = len(rows)
        self.item_slots = item_slots
        need = np.zeros((self.nphase, nslot), dtype=np.int32)
        for q, sl in enumerate(item_slots):
            sl = sl[sl >= 0]
            need[q] = np.bincount(sl, minlength=nslot)[:nslot]
        prev = np.full((self.nphase, nslot), -1, dtype=np.int32)
        last = np.full(nslot, -1, dtype=np.int32)
        for q in range(self.nphase):
            prev[q] = last
            last = np.where(need[q] > 0, q, last)
        self.prog = torch.from_numpy(np.array(rows, dtype=L.PHASE_DTYPE).view(np.uint8)
                                     .copy()).to(dev)
        self.need = torch.from_numpy(need).to(dev)
        self.prev = torch.from_numpy(prev).to(dev)
        self.cnt = torch.zeros((self.nphase, nslot), dtype=torch.int32, device=dev)
        self.fail = torch.zeros(1, dtype=torch.int32, device=dev)
        self.ctl = L.PersistCtl(self.cnt.data_ptr(), self.need.data_ptr(), self.prev.data_ptr(),
                                self.fail.data_ptr(), nslot, 0)
        self.items = int(sum(int(r["nitems"]) for r in rows))

    def launch(self, s, workgroups: int, timeout_s: float = 2.0):
        """Run the program on stream ``s`` (the counters and the failure word
        are cleared first, with fill kernels)."""
        self.cnt.zero_()
        self.fail.zero_()
        pool = self.pool
        L.call("hfm_mg_persistent", self.prog.data_ptr(), self.nphase, L.C.byref(self.ctl),
               pool.sc.data_ptr(), pool.store_id, pool.outer_id, int(workgroups),
               float(timeout_s), s)

    def failed(self) -> bool:
        return int(self.fail.item()) != 0


def persistent_workgroups(pool: "Pool") -> int:
    """Resident workgroups of the persistent kernel for the pool's storage types."""
    n = L.C.c_int32()
    L.call("hfm_mg_persistent_workgroups", pool.store_id, pool.outer_id, L.C.byref(n))
    return int(n.value)


class MGBatch:
    """A fixed batch of instances of given shapes sharing every launch.

    ``set_problem`` (or ``stage_sigma`` + ``build_faces``) loads an instance,
    ``setup`` builds the hierarchies, ``solve`` runs FCG for all of them.
    On a CPU device only the work tables are built (host-side checks).
    """

    def __init__(self, shapes: Sequence[Sequence[int]], storage: str = "bf16", tol: float = 1e-8,
                 maxit: int = 200, nu1: int = 2, nu2: int = 2, omega: float = 2.0 / 3.0,
                 alpha: float = 1.6, max_coarse_edge: int = 4, device=None,
                 iters_per_graph: int = 4, use_graph: bool = True, shrink: float = 0.5,
                 outer: Optional[str] = None, blocks: Optional[dict] = None,
                 coarse_cells: int = 4096, fuse: bool = True, fuse_levels: int = 2,
                 fuse_min_tiles: int = 256, streams: int = 2, method: str = "fcg",
                 tile_planes: int = 32, alloc=None):
        self.params = Params(storage, tol, maxit, nu1, nu2, omega, alpha, max_coarse_edge, outer,
                             blocks, coarse_cells, fuse, fuse_levels, fuse_min_tiles, method,
                             tile_planes)
        shapes = [check_shape(s, max_coarse_edge)[0] for s in shapes]
        if not shapes:
            raise ValueError("empty batch")
        self.pool = Pool(len(shapes), sum(fine_blocks(s) for s in shapes), self.params, device,
                         alloc=alloc)
        self._set_fusion(shapes)
        self.device = self.pool.device
        self.plan_only = self.pool.plan_only
        self.storage = storage
        self.shapes = shapes
        self.insts = [Instance(self.pool, s) for s in shapes]
        self.ninst = len(shapes)
        self.iters_per_graph = int(iters_per_graph)
        self.use_graph = bool(use_graph)
        self.setup_plan = SetupPlan(self.pool, self.insts)
        self.plan = Plan(self.pool, self.insts)
        self.shrink = float(shrink)
        self._subplans = {}
        self.last_info = None
        self.set_streams(streams)

    def set_streams(self, n: int):
        """Solve the batch as ``n`` (1 or 2) groups of instances, each with its
        own plan on its own stream.  The groups' launches then run beside each
        other, so one group's compute-bound V-cycle passes can overlap the
        other's bandwidth-bound FCG passes; results are bitwise the same
        (MI355X, 512-instance batch: 2420 vs 2393 solves/s)."""
        if n not in (1, 2):
            raise ValueError("streams must be 1 or 2")
        self.streams = n
        self._groups = None
        if n == 2 and len(self.insts) >= 2:
            cells = np.cumsum([i.N for i in self.insts])
            h = int(np.searchsorted(cells, cells[-1] / 2)) + 1
            h = min(max(h, 1), len(self.insts) - 1)
            groups = [self.insts[:h], self.insts[h:]]
            self._groups = [(g, (min(i.slot for i in g), max(i.slot for i in g) + 1))
                            for g in groups]
            lo0, hi0 = self._groups[0][1]
            lo1, hi1 = self._groups[1][1]
            if not (hi0 <= lo1 or hi1 <= lo0):     # slots of the groups interleave
                self._groups = None

    def _need_gpu(self):
        if self.plan_only:
            raise RuntimeError("this MGBatch was built on the CPU for planning only")

    def stage_sigma(self, i: int, sx, sy, sz, b):
        self.insts[i].stage_sigma(sx, sy, sz, b)

    def set_faces(self, i: int, faces, b):
        self.insts[i].set_faces(faces, b)

    def set_problem(self, i: int, sx, sy, sz, b, inv_h2: float = 1.0):
        """Load instance ``i`` from cell conductivities and rhs and build its faces."""
        self._need_gpu()
        inst = self.insts[i]
        inst.stage_sigma(sx, sy, sz, b)
        N8 = _align(inst.N) * 4
        base = inst.sigma.data_ptr()
        L.call("hfm_mg_faces", base, base + N8, base + 2 * N8, L.ptr_array(inst.faces(0, False)),
               *inst.shape, float(inv_h2), L.stream_handle(self.device))

    def build_faces(self, inv_h2: float = 1.0):
        """Face coefficients of every staged instance (one batched launch)."""
        self._need_gpu()
        self.setup_plan.build_faces(L.stream_handle(self.device), inv_h2)

    def setup(self):
        """Galerkin coarse operators, storage copies and coarsest inverses."""
        self._need_gpu()
        self.setup_plan.build_hierarchy(L.stream_handle(self.device))

    def restart(self):
        """(Re)initialise every instance: x = 0, r = b."""
        self.pool.stage(self.insts)
        self.plan.init(L.stream_handle(self.device))

    def set_blocks(self, blocks: Optional[dict]):
        """Use these per-launch block sizes (see ``hfm.tune``); rebuilds the plans."""
        for k, b in (blocks or {}).items():
            if b not in BLOCKS:
                raise ValueError(f"block size {b} for {k} is not one of {BLOCKS}")
        self.params.blocks = dict(blocks) if blocks else None
        self._replan()

    def set_coarse_cells(self, n: int):
        """Run levels of at most ``n`` cells in the one-launch coarse cycle; rebuilds the plans."""
        if int(n) < 0:
            raise ValueError("coarse_cells must be >= 0")
        self.params.coarse_cells = int(n)
        self._replan()

    def _fuse_fine(self, shapes, level: int = 0) -> bool:
        """Whether the batch fuses the pass pairs of ``level``: at least
        ``fuse_min_tiles`` tiles there over its instances."""
        p = self.params
        if not p.fuse:
            return False
        tiles = []
        for s in shapes:
            ls = level_shapes(s, p.max_coarse_edge)
            if level < len(ls):
                tiles.append(L.tile(ls[level], p.tile_planes))
        return sum(t[2] for t in tiles if t is not None) >= p.fuse_min_tiles

    def _set_fusion(self, shapes):
        self.pool.fuse_fine = self._fuse_fine(shapes)
        self.pool.fuse_coarse = self._fuse_fine(shapes, 1)

    def set_tile_planes(self, n: int):
        """At most ``n`` planes per tile of the fused pass pairs; rebuilds the
        plans (the pairs' outputs do not depend on the tiling; the fine post
        pair's r.z partials are summed per tile)."""
        if int(n) < 2:
            raise ValueError("tile_planes must be >= 2")
        self.params.tile_planes = int(n)
        self._set_fusion(self.shapes)
        for inst in self.insts:
            inst._tpl = None
        self._replan()

    def set_schedule(self, nu1: int, nu2: int, alpha: float):
        """The V-cycle's sweeps before / after the coarse correction and the
        correction's over-relaxation factor; rebuilds the plans."""
        if int(nu1) < 1 or int(nu2) < 1:
            raise ValueError("nu1 and nu2 must be >= 1")
        self.params.nu1, self.params.nu2, self.params.alpha = int(nu1), int(nu2), float(alpha)
        for inst in self.insts:
            inst._tpl = None
        self._replan()

    def set_fuse(self, fuse: bool, levels: Optional[int] = None,
                 min_tiles: Optional[int] = None):
        """Run the fused pass pairs (default) or every pass as its own launch,
        on the finest ``levels`` levels, the fine one when the batch has at
        least ``min_tiles`` tiles there; rebuilds the plans.  All give the same
        V-cycle (the r.z sum in another order when the fine pair changes)."""
        self.params.fuse = bool(fuse)
        if levels is not None:
            self.params.fuse_levels = int(levels)
        if min_tiles is not None:
            self.params.fuse_min_tiles = int(min_tiles)
        self._set_fusion(self.shapes)
        for inst in self.insts:
            inst._tpl = None
        self._replan()

    def _replan(self):
        self.plan = Plan(self.pool, self.insts)
        self._subplans = {}

    def _subplan(self, live, slots=None):
        key = (tuple(i.slot for i in live), slots)
        plan = self._subplans.get(key)
        if plan is None:
            plan = Plan(self.pool, live, slots)
            self._subplans[key] = plan
        return plan

    def _solve_groups(self):
        """``solve`` with the batch split in two groups on two streams."""
        s = L.stream_handle(self.device)
        self.restart()
        act = self.pool.active_view()
        cur = torch.cuda.current_stream(self.device)
        if getattr(self, "_streams", None) is None:
            self._streams = [torch.cuda.Stream(self.device) for _ in self._groups]
        plans = [self._subplan(g, r) for g, r in self._groups]
        its, maxit = 0, self.params.maxit
        m = self.iters_per_graph + self.iters_per_graph % 2
        while its < maxit and any(p is not None for p in plans):
            for plan, st in zip(plans, self._streams):
                if plan is None:
                    continue
                st.wait_stream(cur)
                with torch.cuda.stream(st):
                    if self.use_graph:
                        g, m = plan.graph(m)
                        g.replay()
                    else:
                        hs = L.stream_handle(self.device)
                        for q in range(m):
                            plan.iteration(hs, q % 2)
            for st in self._streams:
                cur.wait_stream(st)
            its += m
            flags = act.cpu().numpy()
            for q, plan in enumerate(plans):
                if plan is None:
                    continue
                live = [i for i in plan.insts if flags[i.slot]]
                if not live:
                    plans[q] = None
                elif len(live) <= self.shrink * len(plan.insts):
                    plans[q] = self._subplan(live, self._groups[q][1])
        info = self.scalars()
        self.last_info = info
        return info

    def solve(self, persistent: bool = False, log=None):
        """Run FCG to ``tol`` (or ``maxit``) for every instance; returns the scalar records.

        Iterations run ``iters_per_graph`` at a time (one HIP-graph replay, or
        with ``persistent`` one hfm_mg_persistent launch of a program of
        those iterations); between them the active flags are read and, once at
        most ``shrink`` of the plan's instances are still iterating, the
        launches continue on a plan of the active ones only (no empty blocks).
        Both modes give bitwise the same results.  ``log``: a
        ``hfm.solvelog.SolveLog`` that receives one JSON line per instance.
        """
        self._need_gpu()
        t0 = time.perf_counter()
        if persistent:
            info = self._solve_persistent()
        elif self.streams == 2 and self._groups is not None:
            info = self._solve_groups()
        else:
            info = self._solve_plan()
        if log is not None:
            from .solvelog import solve_record
            wall = time.perf_counter() - t0
            for q, inst in enumerate(self.insts):
                log.write(solve_record(inst, self.params, info[q], wall, self.ninst,
                                       self.iteration_bytes(q),
                                       mode="persistent" if persistent else "graphs"))
        return info

    def iteration_bytes(self, i: int) -> float:
        """Algorithmic bytes of one iteration of instance ``i`` (iteration_bytes)."""
        return iteration_bytes(self.insts[i], self.params)

    def _solve_plan(self):
        s = L.stream_handle(self.device)
        self.restart()
        act = self.pool.active_view()
        its = 0
        maxit = self.params.maxit
        plan = self.plan
        m = self.iters_per_graph + self.iters_per_graph % 2
        while its < maxit:
            if self.use_graph:
                g, m = plan.graph(m)
                g.replay()
            else:
                for q in range(m):
                    plan.iteration(s, q % 2)
            its += m
            flags = act.cpu().numpy()
            live = [i for i in plan.insts if flags[i.slot]]
            if not live:
                break
            if len(live) <= self.shrink * len(plan.insts):
                plan = self._subplan(live)
        info = self.scalars()
        self.last_info = info
        return info

    def _persistent(self, live) -> PersistentProgram:
        """The persistent program of ``iters_per_graph`` iterations over the
        instances ``live`` (cached), on a plan with 256-thread cell blocks."""
        m = self.iters_per_graph + self.iters_per_graph % 2
        key = ("persistent", tuple(i.slot for i in live), m)
        prog = self._subplans.get(key)
        if prog is None:
            blocks, self.params.blocks = self.params.blocks, None
            try:
                plan = Plan(self.pool, live)
            finally:
                self.params.blocks = blocks
            prog = PersistentProgram(plan, m)
            self._subplans[key] = prog
        return prog

    def _solve_persistent(self):
        s = L.stream_handle(self.device)
        if getattr(self, "_pwg", None) is None:
            self._pwg = persistent_workgroups(self.pool)
        self.restart()
        act = self.pool.active_view()
        its, maxit = 0, self.params.maxit
        live = list(self.insts)
        prog = self._persistent(live)
        while its < maxit:
            prog.launch(s, self._pwg)
            its += prog.m
            flags = act.cpu().numpy()
            if prog.failed():
                raise L.HfmError("hfm_mg_persistent: a dependency wait timed out")
            now = [i for i in live if flags[i.slot]]
            if not now:
                break
            if len(now) <= self.shrink * len(live):
                live = now
                prog = self._persistent(live)
        info = self.scalars()
        self.last_info = info
        return info

    def scalars(self) -> np.ndarray:
        return self.pool.scalars()[[i.slot for i in self.insts]]

    def solution(self, i: int) -> torch.Tensor:
        return self.insts[i].solution()

    def nbytes(self) -> int:
        return sum(i.nbytes() for i in self.insts) + self.pool.sc.numel() + \
            self.pool.dots.numel() * 8

    # measurement hooks
    def launch_update(self, parity: int = 0):
        self.plan.launch_update(parity, L.stream_handle(self.device))

    def set_xlag(self, lag: int):
        """Select the update kernel's variant for timing: 1 = with the solution
        update (every other iteration), 0 = without.  ``restart`` resets it."""
        idx = torch.tensor([i.slot for i in self.insts], device=self.device)
        self.pool.xlag_view().index_fill_(0, idx, int(lag))

    def _fine_pre_op(self):
        ops = [o for o in self.plan.ops_down if o[5] == 0]
        fused = next((o for o in ops if o[0] == "smooth_restrict"), None)
        return fused or next(o for o in ops if o[0] == "sweep" and o[1] == L.SWEEP["plain"])

    def launch_fine_pre(self):
        """The fine level's second pre-smoothing launch: the fused sweep +
        restriction, or the PLAIN sweep when the plan does not fuse."""
        self.plan._launch(self._fine_pre_op(), L.stream_handle(self.device))

    def fine_pre_kernel(self) -> str:
        return "k_smooth_pair SR" if self._fine_pre_op()[0] == "smooth_restrict" else \
            "k_sweep PLAIN"

    def fine_pre_bytes_per_cell(self) -> float:
        """Algorithmic bytes per fine cell of ``launch_fine_pre``: x, b, 3 faces
        read and the sweep written (T), plus the coarse rhs (T / 8) when fused."""
        e = ELEM[self.storage]
        return 6 * e + (e / 8 if self._fine_pre_op()[0] == "smooth_restrict" else 0)

    def update_bytes_per_cell(self, lag: int = 1) -> int:
        """Algorithmic bytes of the fused k_fcg_update per fine cell: p (outer
        type), 3 FP32 faces, r64 read + written, rlow (T) and the first sweep
        (T) written; with ``lag`` (every other iteration) also x64 read +
        written and the previous direction read."""
        eo = ELEM[self.params.outer]
        return 2 * ELEM[self.storage] + eo + 12 + 16 + (16 + eo if lag else 0)


