"""MI355X (gfx950) stencil + multigrid hot path of the AI-HFM-Solver operator surface.

Modules:
  layers  reference-shaped layer callables on (1, D, H, W, 1) tensors
  mg      batched multigrid-preconditioned FP64 CG solver (work-table kernels)
  _lib    ctypes binding of libhfm.so (include/hfm.h)
"""
from ._lib import HfmError, load  # noqa: F401

__all__ = ["HfmError", "load"]
