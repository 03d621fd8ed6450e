"""Batched multigrid-preconditioned FP64 flexible-CG solver for the
anisotropic resistivity forward problem ``div(sigma grad phi) = b``.

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

Everything runs in libhfm.so work-table kernels: one launch per
(phase, level) covers every instance of the batch, so uneven sizes share
launches and an instance's arithmetic is the same alone or in a batch.  The
per-iteration launch sequence is captured into a HIP graph.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L

CHUNK = 2048          # cells per block of the cell kernels (256 threads x 8)
RCHUNK = 1024         # coarse cells per block of the residual + restriction
ALIGN = 8             # element alignment of every per-instance array offset
SCHUNK = 256          # cells per block of the batched setup kernels
ELEM = {"f32": 4, "bf16": 2}
TORCH_STORE = {"f32": torch.float32, "bf16": torch.bfloat16}


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


def _plane_sizes(s):
    D, H, W = s
    return D * H, D * W, H * W


@dataclasses.dataclass
class _Layout:
    """Block layout of one launch geometry: per member (instance) blk0/nblk."""
    members: List[int]
    blk0: List[int]
    nblk: List[int]
    blk_task: torch.Tensor  # uint16 per block (device)
    nblocks: int


class MGBatch:
    """Resident multigrid hierarchies for a batch of instances of given shapes.

    Allocation depends only on the shapes; ``set_problem`` loads the
    conductivities and right-hand side of one instance, ``setup`` builds the
    Galerkin hierarchies, ``solve`` runs the FCG iterations.
    """

    def __init__(self, shapes: Sequence[Sequence[int]], storage: str = "bf16", tol: float = 1e-8,
                 maxit: int = 200, nu1: int = 2, nu2: int = 2, omega: float = 2.0 / 3.0,
                 alpha: float = 1.6, max_coarse_edge: int = 4, device=None,
                 iters_per_graph: int = 4, use_graph: bool = True):
        if storage not in ELEM:
            raise ValueError(f"storage must be one of {sorted(ELEM)}")
        if nu1 < 1 or nu2 < 1:
            raise ValueError("nu1 and nu2 must be >= 1")
        self.device = torch.device(device or "cuda")
        # A CPU device builds the work tables only (host-side checks); solving needs the GPU.
        self.plan_only = self.device.type != "cuda"
        if not self.plan_only:
            if not torch.cuda.is_available():
                raise RuntimeError("MGBatch needs a GPU (the HIP library has no CPU path)")
            L.load()
        self.storage = storage
        self.store_id = L.STORE[storage]
        self.tol, self.maxit = float(tol), int(maxit)
        self.nu1, self.nu2, self.omega, self.alpha = int(nu1), int(nu2), float(omega), float(alpha)
        self.iters_per_graph = int(iters_per_graph)
        self.use_graph = bool(use_graph)
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        if not self.shapes:
            raise ValueError("empty batch")
        if len(self.shapes) > 65535:
            raise ValueError("at most 65535 instances per batch")
        self.levels = [level_shapes(s, max_coarse_edge) for s in self.shapes]
        for s, lv in zip(self.shapes, self.levels):
            if len(s) != 3 or min(s) < 1:
                raise ValueError(f"bad shape {s}")
            if len(lv) < 2:
                raise ValueError(f"grid {s} is too small for a multigrid hierarchy")
            if math.prod(s) >= 2 ** 31:
                raise ValueError(f"grid {s} exceeds 2^31 cells")
            if math.prod(lv[-1]) > 64:
                raise ValueError(f"coarsest level {lv[-1]} exceeds 64 cells")
        self.ninst = len(self.shapes)
        self.nlev = max(len(lv) for lv in self.levels)
        self._allocate()
        self._build_tables()
        self._graphs = {}
        self.last_info = None

    # ------------------------------------------------------------------ memory
    def _allocate(self):
        dev, st = self.device, TORCH_STORE[self.storage]
        n = self.ninst
        self.cell_off = [[None] * self.nlev for _ in range(n)]
        self.face_off = [[None] * self.nlev for _ in range(n)]
        self.lvl_cells = []
        self.face32, self.faceT = [], []
        self.X, self.Y, self.B = [], [], []
        for lev in range(self.nlev):
            ncell = 0
            planes = [0, 0, 0]
            members = [i for i in range(n) if len(self.levels[i]) > lev]
            for i in members:
                s = self.levels[i][lev]
                self.cell_off[i][lev] = ncell
                ncell += _align(math.prod(s))
            # face buffer: [cx | cy | cz | bxh | byh | bzh], each concatenated over members
            for i in members:
                ps = _plane_sizes(self.levels[i][lev])
                self.face_off[i][lev] = [None] * 6
                for q in range(3):
                    self.face_off[i][lev][3 + q] = planes[q]
                    planes[q] += _align(ps[q])
            base = [0, ncell, 2 * ncell, 3 * ncell, 3 * ncell + planes[0],
                    3 * ncell + planes[0] + planes[1]]
            total = 3 * ncell + sum(planes)
            for i in members:
                fo = self.face_off[i][lev]
                co = self.cell_off[i][lev]
                self.face_off[i][lev] = [base[0] + co, base[1] + co, base[2] + co,
                                         base[3] + fo[3], base[4] + fo[4], base[5] + fo[5]]
            self.lvl_cells.append(ncell)
            f32 = torch.zeros(total, dtype=torch.float32, device=dev)
            self.face32.append(f32)
            self.faceT.append(f32 if self.storage == "f32" else
                              torch.zeros(total, dtype=st, device=dev))
            self.X.append(torch.zeros(ncell, dtype=st, device=dev))
            self.Y.append(torch.zeros(ncell, dtype=st, device=dev))
            self.B.append(torch.zeros(ncell, dtype=st, device=dev))
        N0 = self.lvl_cells[0]
        self.sigma = torch.zeros(3 * N0, dtype=torch.float32, device=dev)  # staged sx | sy | sz
        self.x64 = torch.zeros(N0, dtype=torch.float64, device=dev)
        self.r64 = torch.zeros(N0, dtype=torch.float64, device=dev)
        self.b32 = torch.zeros(N0, dtype=torch.float32, device=dev)
        self.P = [torch.zeros(N0, dtype=st, device=dev), torch.zeros(N0, dtype=st, device=dev)]
        self.ainv_off = []
        tot = 0
        for i in range(n):
            nc = math.prod(self.levels[i][-1])
            self.ainv_off.append(tot)
            tot += nc * nc
        self.ainv = torch.zeros(tot, dtype=torch.float64, device=dev)
        self.sc_np = np.zeros(n, dtype=L.SCALARS_DTYPE)
        self.sc = torch.zeros(n * L.SCALARS_DTYPE.itemsize, dtype=torch.uint8, device=dev)

    def nbytes(self) -> int:
        """Device bytes held by this batch."""
        ts = [*self.face32, *self.X, *self.Y, *self.B, self.x64, self.r64, self.b32, *self.P,
              self.ainv, self.sc, self.dots, self.sigma]
        if self.storage != "f32":
            ts += self.faceT
        return sum(t.numel() * t.element_size() for t in ts)

    # ------------------------------------------------------------------ addresses
    def _faces(self, i, lev, low: bool):
        buf = self.faceT[lev] if low else self.face32[lev]
        es = buf.element_size()
        return [buf.data_ptr() + o * es for o in self.face_off[i][lev]]

    def _vec(self, buf, i, lev):
        return buf.data_ptr() + self.cell_off[i][lev] * buf.element_size()

    def _fine(self, buf, i):
        return buf.data_ptr() + self.cell_off[i][0] * buf.element_size()

    # ------------------------------------------------------------------ tables
    def _layout(self, members, cells_of, chunk=CHUNK) -> _Layout:
        blk0, nblk, bt = [], [], []
        g = 0
        for idx, i in enumerate(members):
            nb = _round8((cells_of(i) + chunk - 1) // chunk)
            blk0.append(g)
            nblk.append(nb)
            bt.append(np.full(nb, idx, dtype=np.uint16))
            g += nb
        arr = np.concatenate(bt) if bt else np.zeros(0, np.uint16)
        t = torch.from_numpy(arr.view(np.int16).copy()).to(self.device)
        return _Layout(members, blk0, nblk, t, g)

    def _task(self, i, lev, lay: _Layout, idx: int, coarse=None):
        t = np.zeros((), dtype=L.TASK_DTYPE)
        D, H, W = self.levels[i][lev]
        t["D"], t["H"], t["W"], t["inst"] = D, H, W, i
        t["blk0"], t["nblk"], t["slot0"] = lay.blk0[idx], lay.nblk[idx], lay.blk0[idx]
        t["magW"], t["magH"] = L.magic(W), L.magic(H)
        if coarse is not None:
            Dc, Hc, Wc = coarse
            t["Dc"], t["Hc"], t["Wc"] = Dc, Hc, Wc
            t["magWc"], t["magHc"] = L.magic(Wc), L.magic(Hc)
        return t

    def _upload(self, tasks):
        arr = np.array(tasks, dtype=L.TASK_DTYPE)
        return torch.from_numpy(arr.view(np.uint8).copy()).to(self.device)

    def _build_tables(self):
        n, nlev = self.ninst, self.nlev
        lv = self.levels
        # post result buffer at a non-coarsest level: pre ends in buf(nu1-1),
        # post sweep s writes the buffer other than the one it reads.
        pre_buf = [self.X, self.Y][(self.nu1 - 1) % 2]
        post_final = [self.X, self.Y][(self.nu1 + self.nu2 - 1) % 2]
        self._pre_buf_is_x = (self.nu1 - 1) % 2 == 0

        def E(i, lev):  # buffer holding the level's correction after its cycle
            if lev == len(lv[i]) - 1:
                return self._vec(self.X[lev], i, lev)
            return self._vec(post_final[lev], i, lev)

        self.ops_down, self.ops_up = [], []
        for lev in range(nlev - 1):
            members = [i for i in range(n) if len(lv[i]) - 1 > lev]
            if not members:
                continue
            lay = self._layout(members, lambda i: math.prod(lv[i][lev]))
            # pre sweeps
            for s in range(self.nu1):
                src = [self.X, self.Y][(s - 1) % 2] if s else None
                dst = [self.X, self.Y][s % 2]
                tasks = []
                for idx, i in enumerate(members):
                    t = self._task(i, lev, lay, idx)
                    t["c"] = self._faces(i, lev, True)
                    t["xin"] = self._vec(src[lev], i, lev) if src else 0
                    t["xout"] = self._vec(dst[lev], i, lev)
                    t["b"] = self._vec(self.B[lev], i, lev)
                    tasks.append(t)
                self.ops_down.append(("sweep", L.SWEEP["zero" if s == 0 else "plain"],
                                      self._upload(tasks), lay, False))
            # residual + restriction to lev+1
            rlay = self._layout(members, lambda i: math.prod(lv[i][lev + 1]), RCHUNK)
            tasks = []
            for idx, i in enumerate(members):
                t = self._task(i, lev, rlay, idx, coarse=lv[i][lev + 1])
                t["c"] = self._faces(i, lev, True)
                t["xin"] = self._vec(pre_buf[lev], i, lev)
                t["b"] = self._vec(self.B[lev], i, lev)
                t["xout"] = self._vec(self.B[lev + 1], i, lev + 1)
                tasks.append(t)
            self.ops_down.append(("restrict", 0, self._upload(tasks), rlay, False))
            # post sweeps (executed in reverse level order)
            ups = []
            for s in range(self.nu2):
                k = self.nu1 + s
                src = [self.X, self.Y][(k - 1) % 2]
                dst = [self.X, self.Y][k % 2]
                tasks = []
                for idx, i in enumerate(members):
                    t = self._task(i, lev, lay, idx, coarse=lv[i][lev + 1])
                    t["c"] = self._faces(i, lev, True)
                    t["xin"] = self._vec(src[lev], i, lev)
                    t["xout"] = self._vec(dst[lev], i, lev)
                    t["b"] = self._vec(self.B[lev], i, lev)
                    if s == 0:
                        t["ec"] = E(i, lev + 1)
                    tasks.append(t)
                dot = lev == 0 and s == self.nu2 - 1
                ups.append(("sweep", L.SWEEP["correct" if s == 0 else "plain"],
                            self._upload(tasks), lay, dot))
            self.ops_up.insert(0, ups)
        # coarsest solves (one task per instance)
        tasks = []
        for i in range(n):
            lc = len(lv[i]) - 1
            t = np.zeros((), dtype=L.TASK_DTYPE)
            D, H, W = lv[i][lc]
            t["D"], t["H"], t["W"], t["inst"] = D, H, W, i
            t["magW"], t["magH"] = L.magic(W), L.magic(H)
            t["c"] = self._faces(i, lc, False)  # FP32 faces for the inverse
            t["b"] = self._vec(self.B[lc], i, lc)
            t["xout"] = self._vec(self.X[lc], i, lc)
            t["ec"] = self.ainv.data_ptr() + self.ainv_off[i] * 8
            tasks.append(t)
        self.coarse_tasks = self._upload(tasks)
        inv_tasks = []
        for i, t in enumerate(tasks):
            t2 = t.copy()
            t2["xout"] = self.ainv.data_ptr() + self.ainv_off[i] * 8
            inv_tasks.append(t2)
        self.inverse_tasks = self._upload(inv_tasks)
        # fine-level FCG tables
        flay = self._layout(list(range(n)), lambda i: math.prod(lv[i][0]))
        self.flay = flay
        self.dots = torch.zeros(max(flay.nblocks, 1), dtype=torch.float64, device=self.device)
        z = post_final[0]
        base = []
        for idx, i in enumerate(range(n)):
            t = self._task(i, 0, flay, idx)
            t["c"] = self._faces(i, 0, False)
            t["x64"] = self._fine(self.x64, i)
            t["r64"] = self._fine(self.r64, i)
            t["rlow"] = self._vec(self.B[0], i, 0)
            t["b32"] = self._fine(self.b32, i)
            t["b"] = self._vec(z[0], i, 0)
            base.append(t)
        self.init_tasks = self._upload(base)
        self.dir_tasks, self.upd_tasks = [], []
        for par in range(2):
            dt_, ut_ = [], []
            for i, t in enumerate(base):
                d = t.copy()
                d["xin"] = self._fine(self.P[par], i)
                d["xout"] = self._fine(self.P[1 - par], i)
                dt_.append(d)
                u = t.copy()
                u["xin"] = self._fine(self.P[1 - par], i)
                ut_.append(u)
            self.dir_tasks.append(self._upload(dt_))
            self.upd_tasks.append(self._upload(ut_))
        for i in range(n):
            self.sc_np[i]["slot0"] = flay.blk0[i]
            self.sc_np[i]["nslot"] = flay.nblk[i]
        # setup kernels: faces from the staged conductivities, Galerkin per level
        slay = self._layout(list(range(n)), lambda i: math.prod(lv[i][0]), SCHUNK)
        N0, es = self.lvl_cells[0], 4
        tasks = []
        for idx, i in enumerate(range(n)):
            t = self._task(i, 0, slay, idx)
            base = self.sigma.data_ptr() + self.cell_off[i][0] * es
            t["xin"], t["b"], t["ec"] = base, base + N0 * es, base + 2 * N0 * es
            t["o"] = self._faces(i, 0, False)
            tasks.append(t)
        self.faces_tasks, self.faces_layout = self._upload(tasks), slay
        self.galerkin_tables = []
        for lev in range(nlev - 1):
            members = [i for i in range(n) if len(lv[i]) > lev + 1]
            glay = self._layout(members, lambda i: math.prod(lv[i][lev + 1]), SCHUNK)
            tasks = []
            for idx, i in enumerate(members):
                t = self._task(i, lev, glay, idx)
                t["c"] = self._faces(i, lev, False)
                t["o"] = self._faces(i, lev + 1, False)
                tasks.append(t)
            self.galerkin_tables.append((self._upload(tasks), glay))

    # ------------------------------------------------------------------ problem setup
    def set_problem(self, i: int, sx, sy, sz, b, inv_h2: float = 1.0):
        """Load instance ``i``: cell conductivities (D,H,W) and right-hand side (D,H,W)."""
        self._need_gpu()
        shape = self.shapes[i]
        ts = []
        for a in (sx, sy, sz, b):
            a = torch.as_tensor(a).reshape(shape).to(device=self.device, dtype=torch.float32)
            ts.append(a.contiguous())
        sx, sy, sz, b = ts
        s = L.stream_handle(self.device)
        out = L.ptr_array(self._faces(i, 0, False))
        L.call("hfm_mg_faces", sx.data_ptr(), sy.data_ptr(), sz.data_ptr(), out, *shape,
               float(inv_h2), s)
        N = math.prod(shape)
        # stream-ordered: the caching allocator reuses sx.. only after these launches
        self.b32[self.cell_off[i][0]:self.cell_off[i][0] + N].copy_(b.reshape(-1))

    def stage_sigma(self, i: int, sx, sy, sz, b):
        """Keep instance ``i``'s cell conductivities resident (face coefficients are
        built from them by ``build_faces``) and load its right-hand side."""
        shape = self.shapes[i]
        N = math.prod(shape)
        o = self.cell_off[i][0]
        for q, a in enumerate((sx, sy, sz)):
            a = torch.as_tensor(a).to(device=self.device, dtype=torch.float32).reshape(-1)
            if a.numel() != N:
                raise ValueError("conductivity size mismatch")
            self.sigma[q * self.lvl_cells[0] + o:q * self.lvl_cells[0] + o + N].copy_(a)
        b = torch.as_tensor(b).to(device=self.device, dtype=torch.float32).reshape(-1)
        if b.numel() != N:
            raise ValueError("rhs size mismatch")
        self.b32[o:o + N].copy_(b)

    def build_faces(self, inv_h2: float = 1.0):
        """Harmonic face coefficients of every staged instance (one batched launch)."""
        self._need_gpu()
        lay = self.faces_layout
        L.call("hfm_mg_faces_batched", self.faces_tasks.data_ptr(), lay.blk_task.data_ptr(),
               lay.nblocks, float(inv_h2), L.stream_handle(self.device))

    def set_faces(self, i: int, faces, b):
        """Load instance ``i`` from FP32 face arrays (cx, cy, cz, bxh, byh, bzh) and rhs."""
        D, H, W = self.shapes[i]
        shapes = [(D, H, W)] * 3 + [(D, H), (D, W), (H, W)]
        lev0 = self.face32[0]
        for q, (a, shp) in enumerate(zip(faces, shapes)):
            a = torch.as_tensor(a).to(device=self.device, dtype=torch.float32).reshape(-1)
            if a.numel() != math.prod(shp):
                raise ValueError(f"face array {q} has {a.numel()} values, expected {shp}")
            o = self.face_off[i][0][q]
            lev0[o:o + a.numel()].copy_(a)
        b = torch.as_tensor(b).to(device=self.device, dtype=torch.float32).reshape(-1)
        N = D * H * W
        if b.numel() != N:
            raise ValueError("rhs size mismatch")
        self.b32[self.cell_off[i][0]:self.cell_off[i][0] + N].copy_(b)

    def _need_gpu(self):
        if self.plan_only:
            raise RuntimeError("this MGBatch was built on the CPU for planning only")

    def setup(self):
        """Galerkin coarse operators, storage copies and coarsest inverses."""
        self._need_gpu()
        s = L.stream_handle(self.device)
        for tasks, lay in self.galerkin_tables:
            L.call("hfm_mg_galerkin_batched", tasks.data_ptr(), lay.blk_task.data_ptr(),
                   lay.nblocks, s)
        if self.storage != "f32":
            for lev in range(self.nlev):
                L.call("hfm_mg_convert", self.face32[lev].data_ptr(), self.faceT[lev].data_ptr(),
                       self.face32[lev].numel(), self.store_id, s)
        L.call("hfm_mg_coarse_inverse", self.inverse_tasks.data_ptr(), self.ninst, s)

    # ------------------------------------------------------------------ launches
    def _launch(self, op, s, parity=0):
        kind, mode, tasks, lay, dot = op
        scp = self.sc.data_ptr()
        if kind == "sweep":
            L.call("hfm_mg_sweep", tasks.data_ptr(), lay.blk_task.data_ptr(), lay.nblocks, scp,
                   self.store_id, mode, self.omega, self.alpha,
                   self.dots.data_ptr() if dot else 0, s)
        elif kind == "restrict":
            L.call("hfm_mg_resid_restrict", tasks.data_ptr(), lay.blk_task.data_ptr(),
                   lay.nblocks, scp, self.store_id, s)

    def _fin(self, kind, s):
        L.call("hfm_fcg_finalize", self.sc.data_ptr(), self.ninst, self.dots.data_ptr(),
               L.FIN[kind], s)

    def precondition(self, s):
        """z = V(rlow): one V-cycle over the whole batch (launch sequence)."""
        for op in self.ops_down:
            self._launch(op, s)
        L.call("hfm_mg_coarse_solve", self.coarse_tasks.data_ptr(), self.ninst, self.sc.data_ptr(),
               self.store_id, s)
        for ups in self.ops_up:
            for op in ups:
                self._launch(op, s)

    def _iteration(self, s, parity):
        fl = self.flay
        scp, dp = self.sc.data_ptr(), self.dots.data_ptr()
        self.precondition(s)
        self._fin("beta", s)
        L.call("hfm_fcg_dir", self.dir_tasks[parity].data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, scp, self.store_id, dp, s)
        self._fin("alpha", s)
        L.call("hfm_fcg_update", self.upd_tasks[parity].data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, scp, self.store_id, dp, s)
        self._fin("check", s)

    def _start(self, s):
        sc = self.sc_np.copy()
        sc["tol"] = self.tol
        sc["maxit"] = self.maxit
        sc["active"] = 1
        sc["status"] = 0
        self.sc.copy_(torch.from_numpy(sc.view(np.uint8).copy()))
        fl = self.flay
        L.call("hfm_fcg_init", self.init_tasks.data_ptr(), fl.blk_task.data_ptr(), fl.nblocks,
               self.sc.data_ptr(), self.store_id, self.dots.data_ptr(), s)
        self._fin("init", s)

    def _graph(self):
        m = self.iters_per_graph + (self.iters_per_graph % 2)
        if m in self._graphs:
            return self._graphs[m], m
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            s = L.stream_handle(self.device)
            for it in range(m):
                self._iteration(s, it % 2)
        torch.cuda.current_stream(self.device).wait_stream(side)
        self._graphs[m] = g
        return g, m

    # ------------------------------------------------------------------ measurement hooks
    def restart(self):
        """Re-initialise the solver state (x = 0, every instance active)."""
        self._start(L.stream_handle(self.device))

    def launch_update(self, parity: int = 0):
        fl = self.flay
        L.call("hfm_fcg_update", self.upd_tasks[parity].data_ptr(), fl.blk_task.data_ptr(),
               fl.nblocks, self.sc.data_ptr(), self.store_id, self.dots.data_ptr(),
               L.stream_handle(self.device))

    def launch_fine_sweep(self):
        """The fine-level PLAIN (second pre-) sweep of the V-cycle."""
        op = next(o for o in self.ops_down if o[0] == "sweep" and o[1] == L.SWEEP["plain"])
        self._launch(op, L.stream_handle(self.device))

    def update_bytes_per_cell(self) -> int:
        """Algorithmic bytes of k_fcg_update per fine cell: p (T), 3 FP32 faces,
        x64 and r64 read + written, rlow (T) written."""
        return 2 * ELEM[self.storage] + 12 + 32

    def sweep_bytes_per_cell(self) -> int:
        """Algorithmic bytes of a PLAIN sweep per cell: x, b, 3 faces, out (all T)."""
        return 6 * ELEM[self.storage]

    def scalars(self) -> np.ndarray:
        return self.sc.cpu().numpy().view(L.SCALARS_DTYPE).copy()

    def solve(self):
        """Run FCG to ``tol`` (or ``maxit``) for every instance; returns the scalar records."""
        self._need_gpu()
        s = L.stream_handle(self.device)
        self._start(s)
        act = self.sc.view(torch.int32).view(self.ninst, -1)[:, 18]
        its = 0
        if self.use_graph:
            g, m = self._graph()
            while its < self.maxit:
                g.replay()
                its += m
                if not bool(act.any()):
                    break
        else:
            while its < self.maxit:
                self._iteration(s, its % 2)
                its += 1
                if its % self.iters_per_graph == 0 and not bool(act.any()):
                    break
        info = self.scalars()
        self.last_info = info
        return info

    def solution(self, i: int) -> torch.Tensor:
        """FP64 solution of instance ``i`` as a (D,H,W) view."""
        N = math.prod(self.shapes[i])
        o = self.cell_off[i][0]
        return self.x64[o:o + N].view(self.shapes[i])
