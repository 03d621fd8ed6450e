"""Autotuner for the batched multigrid plans: launch shapes, tiles, schedules.

Timed on the batch's own data with HIP events:

* the level size below which an instance's V-cycle tail runs in one
  workgroup-per-instance launch (``hfm_mg_coarse_cycle``) instead of one
  launch per pass and level (``tune_coarse``, whole V-cycles timed);
* the threads per block of the remaining per-level launches (``tune_blocks``);
* the planes per tile of the fused pass pairs (``tune_tiles``, whole V-cycles
  timed): more planes walk fewer halo planes, fewer planes give more blocks;
* the sweep schedule (nu1, nu2) and the coarse-correction factor alpha
  (``tune_schedule``, the objective being the time of the whole solve to its
  tolerance: a schedule trades iterations against the cost of one).

The sweeps without a reduction (every V-cycle sweep but the fine level's
last, which carries the r.z partials) and the residual + restriction launches
accept 128, 256 or 512 threads per block; a block covers 8 (sweeps) or 4
(restriction, coarse) cells per thread, so the choice changes how many cells
each block streams and how many blocks each XCD schedules, not what a cell
computes: results are bitwise identical for every choice.  ``tune_blocks``
times each candidate for every such launch of a batch's plan with HIP events
(all instances active, on the batch's own data) and keeps the fastest;
``TuneCache`` stores the choices per device, storage type and batch size
class in a JSON file so a later batch of the same class starts tuned.

The reference has no tuner (its layers are single Keras convolutions); this
is the launch-shape half of the batched solver's tuning.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Optional, Tuple

import torch

from . import _lib as L
from .mg import BLOCK, BLOCKS, MGBatch, Plan

Key = Tuple


def _graph_time_ms(mg: MGBatch, fn, reps: int) -> float:
    """Average time of ``fn(stream)`` replayed from a HIP graph of ``reps`` calls
    (how the solver runs it: no host launch gaps)."""
    dev = mg.device
    side = torch.cuda.Stream(dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        s = L.stream_handle(dev)
        for _ in range(reps):
            fn(s)
    torch.cuda.current_stream(dev).wait_stream(side)
    g.replay()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _activate(mg: MGBatch):
    """Re-initialise the batch and mark every instance active: an instance
    whose rhs is zero (e.g. a solver set up before its first rhs) would
    otherwise be finished at once and its kernels exit without work.  The
    timed kernels do the same work whatever the data; ``solve`` re-stages."""
    mg.restart()
    idx = torch.tensor([i.slot for i in mg.insts], device=mg.device)
    mg.pool.active_view().index_fill_(0, idx, 1)


def _op_key(op) -> Optional[Key]:
    kind, mode, _, _, dot, lev, _ = op
    if dot or kind not in ("sweep", "restrict"):   # fused pass pairs have a fixed shape
        return None
    if kind == "restrict":
        return ("restrict", lev)
    return ("sweep", {v: k for k, v in L.SWEEP.items()}[mode], lev)


def tune_blocks(mg: MGBatch, reps: int = 5, candidates=BLOCKS, min_ms: float = 0.02) -> Dict[Key, int]:
    """Choose the threads per block of every tunable launch of ``mg``'s plan.

    Every candidate plan replays each op ``reps`` times from a HIP graph after
    one warm replay; ops faster than ``min_ms`` with the default keep the default (their
    time is launch latency, not streaming).  The batch's data must be set up
    (``setup``); the solver state is re-initialised first and every instance
    marked active, the launches only rewrite V-cycle scratch vectors, and
    ``restart``/``solve`` re-initialise it again.  Returns the choice table and
    installs it with ``mg.set_blocks``.
    """
    mg._need_gpu()
    _activate(mg)
    times: Dict[Key, Dict[int, float]] = {}
    for b in candidates:
        trial = {}
        base = Plan(mg.pool, mg.insts)           # default table, for the key list
        for op in base.ops_down + [o for ups in base.ops_up for o in ups]:
            k = _op_key(op)
            if k is not None:
                trial[k] = b
        mg.params.blocks = trial
        plan = Plan(mg.pool, mg.insts)
        for op in plan.ops_down + [o for ups in plan.ops_up for o in ups]:
            k = _op_key(op)
            if k is None:
                continue
            times.setdefault(k, {})[b] = _graph_time_ms(mg, lambda s_, op=op: plan._launch(op, s_),
                                                        reps)
    choice = {}
    for k, t in times.items():
        best = min(t, key=t.get)
        if t.get(BLOCK, math.inf) < min_ms or t[best] >= t.get(BLOCK, math.inf):
            best = BLOCK
        choice[k] = best
    mg.set_blocks(choice)
    mg.tune_times = times
    return choice


def tune_coarse(mg: MGBatch, reps: int = 5, candidates=None, max_cells: int = 1 << 17) -> int:
    """Choose the size below which a level joins the one-launch coarse cycle by
    timing whole V-cycles (all instances active, replayed from a HIP graph as
    the solver runs them) for each candidate: by
    default every distinct level size of the batch below the finest, up to
    ``max_cells`` (one workgroup streams such a level in a few microseconds)."""
    mg._need_gpu()
    _activate(mg)
    if candidates is None:
        sizes = {math.prod(lv) for i in mg.insts for lv in i.levels[1:]}
        candidates = sorted(n for n in sizes if n <= max_cells)
    t = {}
    for n in candidates:
        mg.params.coarse_cells = int(n)
        plan = Plan(mg.pool, mg.insts)
        t[n] = _graph_time_ms(mg, lambda s_: plan.precondition(s_), reps)
    best = min(t, key=t.get)
    mg.set_coarse_cells(best)
    mg.coarse_times = t
    return best


TILE_PLANES = (8, 16, 32, 64)
SCHEDULES = ((1, 1, 1.6), (1, 2, 1.6), (2, 1, 1.6), (2, 2, 1.0), (2, 2, 1.3), (2, 2, 1.6),
             (2, 3, 1.6), (3, 3, 1.6))


def tune_tiles(mg: MGBatch, reps: int = 5, candidates=TILE_PLANES) -> int:
    """Choose the planes per tile of the fused pass pairs by timing whole
    V-cycles (all instances active, graph-replayed).  The pairs compute the
    same values for every tiling; only the fine post pair's r.z partials are
    summed per tile."""
    mg._need_gpu()
    t = {}
    for z in candidates:
        mg.set_tile_planes(int(z))
        _activate(mg)
        plan = mg.plan
        t[int(z)] = _graph_time_ms(mg, lambda s_: plan.precondition(s_), reps)
    best = min(t, key=t.get)
    mg.set_tile_planes(best)
    mg.tile_times = t
    return best


def tune_schedule(mg: MGBatch, candidates=SCHEDULES) -> Tuple[int, int, float]:
    """Choose (nu1, nu2, alpha) by the time of the batch's whole solve to its
    tolerance (a warm solve captures each schedule's graphs first); schedules
    whose solve leaves an instance unconverged are discarded.  Installs and
    returns the fastest."""
    mg._need_gpu()
    dev = mg.device
    t = {}
    for sch in candidates:
        mg.set_schedule(*sch)
        info = mg.solve()
        if not bool((info["status"] == 1).all()):
            t[tuple(sch)] = math.inf
            continue
        torch.cuda.synchronize(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        info = mg.solve()
        e1.record()
        e1.synchronize()
        t[tuple(sch)] = e0.elapsed_time(e1) if bool((info["status"] == 1).all()) else math.inf
    best = min(t, key=t.get)
    if not math.isfinite(t[best]):
        raise RuntimeError("no schedule converged")
    mg.set_schedule(*best)
    mg.schedule_times = t
    return best


def tune(mg: MGBatch, reps: int = 5, schedule: bool = False) -> dict:
    """The coarse-cycle threshold, the block sizes of the remaining launches,
    the fused pairs' planes per tile and, with ``schedule``, the sweep
    schedule."""
    n = tune_coarse(mg, reps)
    blocks = tune_blocks(mg, reps)
    planes = tune_tiles(mg, reps)
    out = {"coarse_cells": n, "blocks": blocks, "tile_planes": planes}
    if schedule:
        out["schedule"] = list(tune_schedule(mg))
    return out


def batch_class(mg: MGBatch) -> str:
    """Size class of a batch: storage, outer type and log2 of its fine cells."""
    cells = sum(i.N for i in mg.insts)
    return f"{mg.params.storage}/{mg.params.outer}/2^{int(round(math.log2(max(cells, 1))))}"


class TuneCache:
    """Block-size tables keyed by device name and batch class, in a JSON file."""

    def __init__(self, path: Optional[str] = None):
        self.path = path or os.environ.get(
            "HFM_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "hfm", "tune.json"))
        try:
            with open(self.path) as f:
                self.data = json.load(f)
        except (OSError, ValueError):
            self.data = {}

    @staticmethod
    def _dev(mg: MGBatch) -> str:
        return torch.cuda.get_device_name(mg.device)

    def get(self, mg: MGBatch) -> Optional[dict]:
        rec = self.data.get(self._dev(mg), {}).get(batch_class(mg))
        if rec is None:
            return None
        out = {"coarse_cells": int(rec["coarse_cells"]),
               "blocks": {tuple(json.loads(k)): int(v) for k, v in rec["blocks"].items()}}
        if "tile_planes" in rec:
            out["tile_planes"] = int(rec["tile_planes"])
        if "schedule" in rec:
            nu1, nu2, alpha = rec["schedule"]
            out["schedule"] = [int(nu1), int(nu2), float(alpha)]
        return out

    def put(self, mg: MGBatch, choice: dict):
        rec = {"coarse_cells": int(choice["coarse_cells"]),
               "blocks": {json.dumps(list(k)): int(v) for k, v in choice["blocks"].items()}}
        if "tile_planes" in choice:
            rec["tile_planes"] = int(choice["tile_planes"])
        if "schedule" in choice:
            rec["schedule"] = [int(choice["schedule"][0]), int(choice["schedule"][1]),
                               float(choice["schedule"][2])]
        self.data.setdefault(self._dev(mg), {})[batch_class(mg)] = rec
        os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def tuned(self, mg: MGBatch, reps: int = 5) -> dict:
        """The cached choice for ``mg``'s class if any (installed), else tune and store."""
        choice = self.get(mg)
        if choice is not None:
            mg.params.coarse_cells = choice["coarse_cells"]
            mg.set_blocks(choice["blocks"])
            if "tile_planes" in choice:
                mg.set_tile_planes(choice["tile_planes"])
            if "schedule" in choice:
                mg.set_schedule(*choice["schedule"])
            return choice
        choice = tune(mg, reps)
        self.put(mg, choice)
        return choice
