"""Write tests/golden/fp64_reference_solves.npz: FP64 reference solutions of two
large anisotropic instances (96^3 and 128^3), sampled at fixed cells.

Each instance is the oracle's ``make_anisotropic_fields(shape, seed, sigma_ln,
n_dipoles)`` draw (numpy PCG64), solved by the oracle's FP64 conjugate
gradients preconditioned by an FP64 V(2,2) cycle (``solve_fp64_pcg``) to a
relative residual of 1e-13 on the FP32 face operator.  Stored per instance:
the indices of 8192 cells drawn from ``default_rng(0)``, the solution there,
the solution's 2-norm and its FP64 relative residual.  The GPU test compares
its solve on the same cells (a sampled relative error).  Run from the
repository root: ``python tests/golden/make_fp64_references.py`` (~2 minutes).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import hfm_oracle as O  # noqa: E402

CASES = [((96, 96, 96), 31, 2.0, 3), ((128, 128, 128), 32, 1.5, 2)]
NSAMPLE = 8192


def main():
    out = {}
    for q, (shape, seed, sig, nd) in enumerate(CASES):
        sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, nd)
        op = O.harmonic_faces(sx, sy, sz).astype("f32")
        x = O.solve_fp64_pcg(op, b.astype(np.float32), tol=1e-13)
        r = b.astype(np.float64) - O.face_apply(op, x)
        idx = np.sort(np.random.default_rng(0).choice(x.size, NSAMPLE, replace=False))
        out[f"case{q}_spec"] = np.array([*shape, seed, sig, nd], dtype=np.float64)
        out[f"case{q}_idx"] = idx.astype(np.int64)
        out[f"case{q}_x"] = x.ravel()[idx]
        out[f"case{q}_norm"] = np.array([np.linalg.norm(x)])
        out[f"case{q}_relres"] = np.array([np.linalg.norm(r) / np.linalg.norm(b)])
        print(shape, "relres", out[f"case{q}_relres"][0], flush=True)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fp64_reference_solves.npz"),
             **out)


if __name__ == "__main__":
    main()
