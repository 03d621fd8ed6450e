"""Structured per-solve log: one JSON line per finished solve.

The reference prints a step's "Pressure error / cty equation residual" maxima
(UM/AI_HFM_v3.ipynb:2622-2624) and keeps norm histories as .npy files
(ARTs/advection_diffusion_2D.py:127-149); here every finished solve (of a
batch or of a served request) is one JSON object per line with its shape,
level count, precision, outer method, stop status, iterations (V-cycles),
relative residual, wall time and the algorithmic bytes its iterations moved.
"""
from __future__ import annotations

import io
import json
import threading
from typing import Union

from . import _lib as L


class SolveLog:
    """Appends one JSON line per record to ``sink`` (a path or a text stream)."""

    def __init__(self, sink: Union[str, io.TextIOBase]):
        self._own = isinstance(sink, str)
        self._f = open(sink, "a", encoding="utf-8") if self._own else sink
        self._lock = threading.Lock()

    def write(self, rec: dict) -> None:
        line = json.dumps(rec, sort_keys=True, allow_nan=True)
        with self._lock:
            self._f.write(line + "\n")
            self._f.flush()

    def close(self) -> None:
        if self._own:
            self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def solve_record(inst, params, sc_row, wall_s: float, batch: int, iteration_bytes: float,
                 **extra) -> dict:
    """The record of instance ``inst`` from its scalar record ``sc_row``."""
    iters = max(int(sc_row["iter"]), 0)
    nbytes = float(iteration_bytes) * iters
    rec = {
        "event": "solve",
        "shape": list(inst.shape),
        "levels": len(inst.levels),
        "storage": params.storage,
        "outer": params.outer,
        "method": params.method,
        "status": L.STATUS.get(int(sc_row["status"]), str(int(sc_row["status"]))),
        "iters": iters,
        "rel_residual": float(sc_row["rel"]),
        "tol": float(params.tol),
        "wall_s": round(float(wall_s), 6),
        "batch_instances": int(batch),
        "alg_bytes": nbytes,
        # the batch's instances share the wall time: a solve's bandwidth is its
        # bytes over that time (a lower bound on what its kernels achieved)
        "alg_GBps_share": nbytes / wall_s / 1e9 if wall_s > 0 else None,
    }
    rec.update(extra)
    return rec
