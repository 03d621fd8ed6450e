"""Snapshot / restart of time-step driver fields in the reference's file format.

The reference writes one ``.npy`` file per field every ``n_out`` steps,
named ``<dir>/<field><itime>.npy`` and holding the field's (D, H, W) array
(``values[0, :, :, :, 0]`` of its (1, D, H, W, 1) tensor), and restarts by
loading ``u<n>.npy, v<n>.npy, w<n>.npy, p<n>.npy`` as float32 and continuing at
``ctime = n * dt`` (ARTs/naiver_stokes_3D.py:488-497 ``save_data``, :531-549
the ``Restart`` block; UM/AI_HFM_v3.ipynb cell 19 ``save_data`` and cell 21).

Here the device-to-host copies are asynchronous: ``SnapshotWriter`` copies a
step's fields into pinned host buffers on the stream that produced them and
a writer thread saves them once the copy's event has completed, so the time
loop keeps the GPU busy while files are written.  ``save_data`` is the
reference's synchronous call.  Files hold each field's own dtype (the 3D
drivers' fields are float32, as in the reference; the 2D flow's projection
potential is float64 so that a restart continues bitwise).  Loading never
unpickles (``numpy.load`` with ``allow_pickle=False``).
"""
from __future__ import annotations

import os
import queue
import threading
from typing import Dict, Iterable, List, Mapping, Optional, Tuple

import numpy as np
import torch

REFERENCE_FIELDS = ("u", "v", "w", "p")


def _grid_view(x) -> torch.Tensor:
    """The field's grid: a (1, ..., 1) reference-layout tensor loses its unit axes."""
    t = torch.as_tensor(x)
    if t.dim() in (4, 5) and t.shape[0] == 1 and t.shape[-1] == 1:
        t = t[0, ..., 0]
    return t


def snapshot_path(directory: str, name: str, itime: int) -> str:
    return os.path.join(directory, f"{name}{int(itime)}.npy")


def save_data(fields: Mapping[str, torch.Tensor], n_out: int, itime: int,
              directory: str) -> List[str]:
    """The reference's ``save_data``: when ``itime % n_out == 0`` write every
    field to ``<directory>/<name><itime>.npy`` (blocking); returns the paths."""
    if n_out <= 0 or itime % n_out:
        return []
    os.makedirs(directory, exist_ok=True)
    out = []
    for name, x in fields.items():
        path = snapshot_path(directory, name, itime)
        np.save(path, _grid_view(x).detach().cpu().numpy())
        out.append(path)
    return out


class SnapshotWriter:
    """Asynchronous ``save_data``: pinned host staging + a writer thread.

    ``maybe_save(itime, **fields)`` returns at once; the copies are ordered on
    the current stream of the fields' device after the work that produced
    them.  At most ``depth`` snapshots are staged at a time (the call waits for
    the oldest when all are in flight).  ``close()`` (or leaving a ``with``
    block) waits for every file; errors from the writer thread are raised
    there.
    """

    def __init__(self, directory: str, n_out: int, depth: int = 2):
        if n_out <= 0:
            raise ValueError("n_out must be positive")
        self.directory, self.n_out, self.depth = directory, int(n_out), max(1, int(depth))
        os.makedirs(directory, exist_ok=True)
        self._free: "queue.Queue[Dict[Tuple, torch.Tensor]]" = queue.Queue()
        for _ in range(self.depth):
            self._free.put({})
        self._work: "queue.Queue" = queue.Queue()
        self._error: Optional[BaseException] = None
        self.written: List[str] = []
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    def _run(self):
        while True:
            item = self._work.get()
            if item is None:
                return
            itime, ev, staged, pool = item
            try:
                if ev is not None:
                    ev.synchronize()
                for name, host in staged:
                    path = snapshot_path(self.directory, name, itime)
                    np.save(path, host.numpy())
                    self.written.append(path)
            except BaseException as e:  # reported by close()
                self._error = e
            finally:
                self._free.put(pool)

    def maybe_save(self, itime: int, **fields) -> bool:
        """Stage the fields for ``<name><itime>.npy`` when ``itime % n_out == 0``."""
        if itime % self.n_out:
            return False
        if self._error is not None:
            raise RuntimeError("snapshot writer failed") from self._error
        grids = [(name, _grid_view(x).detach()) for name, x in fields.items()]
        devices = {g.device for _, g in grids if g.is_cuda}
        if len(devices) > 1:
            raise ValueError("snapshot fields must share one device")
        pool = self._free.get()        # waits while `depth` snapshots are in flight
        try:
            staged = []
            for name, g in grids:
                key = (name, tuple(g.shape), g.dtype)
                host = pool.get(key)
                if host is None:
                    host = torch.empty(g.shape, dtype=g.dtype, pin_memory=g.is_cuda)
                    pool[key] = host
                host.copy_(g, non_blocking=g.is_cuda)
                staged.append((name, host))
            ev = None
            if devices:
                ev = torch.cuda.Event()
                ev.record(torch.cuda.current_stream(devices.pop()))
        except BaseException:
            self._free.put(pool)       # the pool goes back: later snapshots must not block
            raise
        self._work.put((int(itime), ev, staged, pool))
        return True

    def close(self):
        if self._thread.is_alive():
            self._work.put(None)
            self._thread.join()
        if self._error is not None:
            raise RuntimeError("snapshot writer failed") from self._error

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def load_restart(directory: str, nrestart: int, dt: float,
                 names: Iterable[str] = REFERENCE_FIELDS, device=None,
                 dtype: Optional[torch.dtype] = None) -> Tuple[Dict[str, torch.Tensor], float]:
    """The reference's restart block: load ``<name><nrestart>.npy`` for every
    name and return ({name: tensor on ``device``}, ctime_old = nrestart * dt).

    ``dtype=None`` keeps each file's dtype; the reference casts to float32
    (``dtype=torch.float32``).  All fields must share one grid shape."""
    out: Dict[str, torch.Tensor] = {}
    missing = [snapshot_path(directory, n, nrestart) for n in names
               if not os.path.exists(snapshot_path(directory, n, nrestart))]
    if missing:
        raise FileNotFoundError(f"restart files missing: {missing}")
    shape = None
    for n in names:
        a = np.load(snapshot_path(directory, n, nrestart), allow_pickle=False)
        if shape is None:
            shape = a.shape
        elif a.shape != shape:
            raise ValueError(f"restart field {n} has shape {a.shape}, expected {shape}")
        t = torch.from_numpy(np.ascontiguousarray(a))
        if dtype is not None:
            t = t.to(dtype)
        out[n] = t.to(device) if device is not None else t
    return out, float(nrestart) * float(dt)
