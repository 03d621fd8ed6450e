"""Continuous-batching request loop for anisotropic forward solves.

Requests (a grid, conductivities and a right-hand side) are queued on
arrival.  Each epoch the loop
  1. admits queued requests while their resident hierarchies fit the HBM
     budget and a solver slot is free (first come, first served),
  2. builds the new hierarchies (one batched setup launch sequence),
  3. rebuilds the launch plan over every resident request when the set
     changed and initialises the newly admitted ones,
  4. runs ``iters_per_epoch`` FCG iterations for all residents in shared
     launches (requests that arrive meanwhile join at the next epoch) and
     queues an asynchronous copy of the solver scalars to pinned memory,
  5. then, while the GPU runs that epoch, reads the previous epoch's scalars
     and retires the requests that had converged (or failed) by then.
A retired request's kernels in the epoch in flight exit at once (its slot
is inactive), so its memory and slot are reused safely by later, stream-
ordered work.  Every request's arithmetic is independent of the others
(per-instance kernels and fixed-order reductions), so a result is bitwise
the same alone and in any mix.  Whether the fine level's pass pairs run
fused (``Params.fuse``, and a capacity of at least FUSE_MIN_CAPACITY slots)
is the server's choice for every plan it builds, so it does not depend on
the mix either.

The reference has no serving layer; this is the request loop around the
multigrid solve of ART/BGS_multigrid_v1.ipynb#c42's operator.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Deque, Dict, List, Optional

import numpy as np
import torch

from . import _lib as L
from .mg import Instance, Params, Plan, Pool, SetupPlan, check_shape, iteration_bytes


@dataclasses.dataclass
class Request:
    rid: int
    shape: tuple
    sx: torch.Tensor
    sy: torch.Tensor
    sz: torch.Tensor
    b: torch.Tensor
    t_arrival: float = 0.0


@dataclasses.dataclass
class Result:
    rid: int
    shape: tuple
    status: str
    iters: int
    rel_residual: float
    x: Optional[torch.Tensor]
    t_arrival: float
    t_admitted: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


FUSE_MIN_CAPACITY = 64     # solver slots from which a server fuses the fine level


class SolveServer:
    """HBM-budgeted continuous batching of multigrid solves on one GPU."""

    def __init__(self, params: Optional[Params] = None, hbm_budget: float = 200e9,
                 capacity: int = 4096, dot_slots: int = 1 << 22, iters_per_epoch: int = 4,
                 device=None, clock=time.perf_counter, log=None):
        """``log``: a ``hfm.solvelog.SolveLog`` receiving one JSON line per
        finished request (its times, iterations, status and residual)."""
        self.params = params or Params()
        self.log = log
        self.pool = Pool(capacity, dot_slots, self.params, device)
        # fused fine-level pass pairs only for a server that batches: with a
        # few residents their tiles leave most CUs idle
        self.pool.fuse_fine = self.params.fuse and capacity >= FUSE_MIN_CAPACITY
        self.pool.fuse_coarse = self.pool.fuse_fine
        if self.pool.plan_only:
            raise RuntimeError("SolveServer needs a GPU")
        self.device = self.pool.device
        self.budget = float(hbm_budget)
        self.used = 0.0
        self.iters_per_epoch = int(iters_per_epoch)
        self.clock = clock
        self.queue: Deque[Request] = collections.deque()
        self.resident: Dict[int, tuple] = {}   # rid -> (request, instance, bytes, t_admitted)
        self.plan: Optional[Plan] = None
        self.parity = 0
        self.epochs = 0
        self.peak_used = 0.0
        nb = self.pool.sc.numel()
        self._host_sc = [torch.empty(nb, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self._inflight = None      # (event, host buffer, residents) of the last launched epoch

    # ------------------------------------------------------------------ intake
    def submit(self, rid: int, sx, sy, sz, b, t_arrival: Optional[float] = None) -> None:
        """Queue a solve of div(sigma grad phi) = b on the grid of ``b``'s shape."""
        shape = check_shape(tuple(b.shape), self.params.max_coarse_edge)[0]
        need = Instance.estimate_bytes(shape, self.params)
        if need > self.budget:
            raise ValueError(f"request {rid}: {need / 1e9:.1f} GB exceeds the HBM budget")
        self.queue.append(Request(rid, shape, sx, sy, sz, b,
                                  self.clock() if t_arrival is None else t_arrival))

    # ------------------------------------------------------------------ epoch
    def _admit(self) -> List[Instance]:
        new = []
        while self.queue and self.pool.free_slots() > 0:
            req = self.queue[0]
            need = Instance.estimate_bytes(req.shape, self.params)
            if self.used + need > self.budget:
                break
            self.queue.popleft()
            inst = Instance(self.pool, req.shape)
            inst.stage_sigma(req.sx, req.sy, req.sz, req.b)
            req.sx = req.sy = req.sz = req.b = None   # the instance holds its own copies
            self.used += need
            self.peak_used = max(self.peak_used, self.used)
            self.resident[req.rid] = (req, inst, need, self.clock())
            new.append(inst)
        return new

    def step(self) -> List[Result]:
        """Launch one epoch, then retire what the previous epoch finished;
        returns those requests."""
        s = L.stream_handle(self.device)
        new = self._admit()
        if new:
            sp = SetupPlan(self.pool, new)
            sp.build_faces(s)
            sp.build_hierarchy(s)
            self.pool.stage(new)
        if new or self.plan is None:
            insts = [v[1] for v in self.resident.values()]
            self.plan = Plan(self.pool, insts) if insts else None
        prev, self._inflight = self._inflight, None
        if self.plan is not None:
            if new:
                self.plan.init(s)
            for _ in range(self.iters_per_epoch):
                self.plan.iteration(s, self.parity)
                self.parity ^= 1
            self.epochs += 1
            buf = self._host_sc[self.epochs % 2]
            buf.copy_(self.pool.sc, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(self.device))
            self._inflight = (ev, buf, dict(self.resident))
        if prev is None:
            return []
        return self._retire(*prev)

    def _retire(self, ev, buf, members) -> List[Result]:
        """Retire the members of an epoch whose scalars (in ``buf``) say they stopped."""
        ev.synchronize()
        sc = buf.numpy().view(L.SCALARS_DTYPE)
        now = self.clock()
        done = []
        for rid, (req, inst, need, t_adm) in members.items():
            if rid not in self.resident:
                continue
            rec = sc[inst.slot]
            if int(rec["active"]) == 1:
                continue
            status = L.STATUS.get(int(rec["status"]), "unknown")
            done.append(Result(rid, req.shape, status, int(rec["iter"]), float(rec["rel"]),
                               inst.solution(), req.t_arrival, t_adm, now))
            if self.log is not None:
                from .solvelog import solve_record
                self.log.write(solve_record(
                    inst, self.params, rec, now - t_adm, len(members),
                    iteration_bytes(inst, self.params), rid=rid,
                    t_arrival=req.t_arrival, t_admitted=t_adm, t_done=now,
                    queue_s=t_adm - req.t_arrival))
            inst.release()
            self.used -= need
            del self.resident[rid]
        if done:
            self.plan = None      # rebuilt over the remaining residents next epoch
        return done

    def pending(self) -> int:
        return len(self.queue) + len(self.resident)

    def drain(self) -> List[Result]:
        out = []
        while self.pending():
            out += self.step()
        return out


def poisson_stream(n: int, rate: float, seed: int, lo: int = 32, hi: int = 256):
    """(t_arrival, shape, seed) for n requests: exponential inter-arrival times
    at ``rate``/s, cube edge log-uniform in [lo, hi] rounded to a multiple of 16."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    edges = 16 * np.maximum(1, np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n)) / 16))
    return [(float(t[q]), (int(e),) * 3, int(seed) * 7919 + q) for q, e in enumerate(edges)]


def run_stream(server: SolveServer, stream, make_fields, clock=time.perf_counter) -> List[Result]:
    """Replay a timed request stream against ``server`` in wall-clock time."""
    t0 = clock()
    results, q = [], 0
    while q < len(stream) or server.pending():
        now = clock() - t0
        while q < len(stream) and stream[q][0] <= now:
            t, shape, seed = stream[q]
            server.submit(q, *make_fields(shape, seed), t_arrival=t0 + t)
            q += 1
        if server.pending():
            results += server.step()
        elif q < len(stream):
            time.sleep(max(0.0, stream[q][0] - (clock() - t0)))
    return results


def latency_stats(results: List[Result]) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": len(results)}
