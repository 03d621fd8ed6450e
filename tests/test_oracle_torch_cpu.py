"""The PyTorch-CPU per-op restatement used as the CPU baseline (CPU only): its
operator is the oracle's face operator, and its FCG solve reaches the FP64
direct solution and the oracle's FP32-storage iteration count within one."""
import numpy as np
import pytest

from oracle import hfm_oracle as O

torch = pytest.importorskip("torch")


def test_operator_is_the_oracle_face_operator():
    from oracle import torch_cpu as TC
    sx, sy, sz, _ = O.make_anisotropic_fields((8, 6, 10), 2, 1.5, 1)
    op = O.harmonic_faces(sx, sy, sz).astype("f32")
    lv = TC._Level(op)
    x = np.random.default_rng(0).standard_normal((8, 6, 10))
    got = TC._apply64(lv, torch.from_numpy(x)).numpy()
    np.testing.assert_allclose(got, O.face_apply(op, x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 24, 24)])
def test_solve_reaches_the_fp64_solution(shape):
    # 24^3 halves to 3^3 like the 48^3 and 96^3 workload edges
    from oracle import torch_cpu as TC
    sx, sy, sz, b = O.make_anisotropic_fields(shape, 3, 1.0, 2)
    x, it, ok, _ = TC.fcg_solve(sx, sy, sz, b, tol=1e-8, threads=2)
    assert ok
    op = O.harmonic_faces(sx, sy, sz).astype("f32")
    xd = O.solve_fp64_direct(op, b.astype(np.float32))
    assert np.linalg.norm(x.reshape(shape) - xd) / np.linalg.norm(xd) < 1e-6
    ref = O.fcg_solve(op, b, tol=1e-8, storage="f32")
    assert abs(it - ref.iters) <= 1
