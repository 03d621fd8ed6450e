"""The oracle against closed forms, library identities and brute force."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import hfm_oracle as O


def test_conv_same_is_torch_cross_correlation():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 6, 7))
    K = rng.standard_normal((3, 3, 3))
    ref = torch.nn.functional.conv3d(torch.from_numpy(x)[None, None], torch.from_numpy(K)[None, None],
                                     padding=1)[0, 0].numpy()
    np.testing.assert_allclose(O.conv_same(x, K), ref, rtol=1e-13, atol=1e-13)
    x2 = rng.standard_normal((6, 7)); K2 = rng.standard_normal((3, 3))
    ref2 = torch.nn.functional.conv2d(torch.from_numpy(x2)[None, None], torch.from_numpy(K2)[None, None],
                                      padding=1)[0, 0].numpy()
    np.testing.assert_allclose(O.conv_same(x2, K2), ref2, rtol=1e-13, atol=1e-13)


def test_a7_const_dirichlet_sine_eigenmodes():
    D, H, W = 6, 7, 9
    Dx, Dy, Dz = 200.0, 100.0, 50.0
    K = O.a7_const_kernel(Dx, Dy, Dz)
    k, j, i = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    for pz, py, px in [(1, 1, 1), (2, 3, 1), (5, 6, 8)]:
        v = (np.sin(np.pi * pz * (k + 1) / (D + 1)) * np.sin(np.pi * py * (j + 1) / (H + 1))
             * np.sin(np.pi * px * (i + 1) / (W + 1)))
        lam = (2 * Dz * (1 - np.cos(np.pi * pz / (D + 1))) + 2 * Dy * (1 - np.cos(np.pi * py / (H + 1)))
               + 2 * Dx * (1 - np.cos(np.pi * px / (W + 1))))
        np.testing.assert_allclose(O.conv_same(v, K), lam * v, atol=1e-10)


def test_restrict_prolong_identities():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6, 8))
    np.testing.assert_allclose(O.restrict_mean(O.prolong_nearest(x)), x, atol=1e-15)
    one = np.ones((8, 8, 8))
    np.testing.assert_array_equal(O.restrict_mean(one), np.ones((4, 4, 4)))
    y = rng.standard_normal((8, 4, 6))
    assert abs(O.restrict_mean(y).mean() - y.mean()) < 1e-14
    # the mean of 8 children, element by element
    assert abs(O.restrict_mean(y)[1, 0, 2] - y[2:4, 0:2, 4:6].mean()) < 1e-15
    np.testing.assert_array_equal(O.prolong_nearest(np.arange(2.0)), [0, 0, 1, 1])


def test_var6_reduces_to_constant_operator_and_brute_force():
    rng = np.random.default_rng(2)
    shape = (4, 5, 6)
    Dx, Dy, Dz = 3.0, 2.0, 1.5
    mats = {"xm1": np.full(shape, Dx), "xp1": np.full(shape, Dx), "ym1": np.full(shape, Dy),
            "yp1": np.full(shape, Dy), "zm1": np.full(shape, Dz), "zp1": np.full(shape, Dz)}
    x = rng.standard_normal(shape)
    np.testing.assert_allclose(O.a_operator_var6(x, mats), O.conv_same(x, O.a7_const_kernel(Dx, Dy, Dz)),
                               atol=1e-12)
    # random fields: dense assembly by unit vectors
    mats = {k: rng.uniform(0.5, 2.0, shape) for k in O.VAR6_DIRS}
    n = int(np.prod(shape))
    A = np.zeros((n, n))
    for q in range(n):
        e = np.zeros(n); e[q] = 1.0
        A[:, q] = O.a_operator_var6(e.reshape(shape), mats).ravel()
    # brute-force row formula for one interior cell
    c = (2, 2, 3)
    row = np.zeros(shape)
    row[c] = sum(mats[k][c] for k in O.VAR6_DIRS)
    for k, off in zip(O.VAR6_DIRS, [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]):
        row[tuple(np.add(c, off))] -= mats[k][c]
    np.testing.assert_allclose(A[np.ravel_multi_index(c, shape)], row.ravel(), atol=1e-14)
    np.testing.assert_allclose(O.var6_tot(mats), sum(mats[k] for k in O.VAR6_DIRS))


def _rand_faces(shape, seed=0):
    rng = np.random.default_rng(seed)
    s = [np.exp(rng.standard_normal(shape)).astype(np.float32) for _ in range(3)]
    return O.harmonic_faces(*s), s


def test_face_operator_matches_sparse_symmetric_and_rowsums():
    op, s = _rand_faces((5, 6, 7))
    A = O.face_sparse(op)
    x = np.random.default_rng(3).standard_normal(op.shape)
    np.testing.assert_allclose(O.face_apply(op, x).ravel(), A @ x.ravel(), rtol=1e-12, atol=1e-12)
    assert abs(A - A.T).max() == 0
    # interior rows sum to zero; boundary rows sum to their boundary-face coefficients
    rs = np.asarray(A.sum(axis=1)).reshape(op.shape)
    assert np.abs(rs[1:-1, 1:-1, 1:-1]).max() < 1e-12
    np.testing.assert_allclose(rs[2, 3, 0], s[0][2, 3, 0], rtol=1e-6)


def test_harmonic_faces_closed_form():
    op, s = _rand_faces((3, 4, 5), seed=4)
    a, b = float(s[0][1, 2, 3]), float(s[0][1, 2, 2])
    assert abs(op.cx[1, 2, 3] - 2 * a * b / (a + b)) < 1e-12
    assert op.bxh[1, 2] == pytest.approx(float(s[0][1, 2, 4]))
    assert op.cz[0, 1, 1] == pytest.approx(float(s[2][0, 1, 1]))


def test_face_operator_uniform_is_a7_sine_modes():
    shape = (5, 6, 7)
    ones = np.ones(shape, np.float32)
    op = O.harmonic_faces(ones * 3.0, ones * 2.0, ones * 1.0)
    x = np.random.default_rng(5).standard_normal(shape)
    np.testing.assert_allclose(O.face_apply(op, x), O.conv_same(x, O.a7_const_kernel(3.0, 2.0, 1.0)),
                               atol=1e-12)


@pytest.mark.parametrize("shape", [(4, 4, 4), (5, 3, 7), (8, 6, 2), (1, 5, 5)])
def test_galerkin_is_ptap_over_8(shape):
    op, _ = _rand_faces(shape, seed=6)
    A = O.face_sparse(op).toarray()
    cs = O.coarse_shape(shape)
    n, nc = int(np.prod(shape)), int(np.prod(cs))
    P = np.zeros((n, nc))
    for idx in np.ndindex(*shape):
        P[np.ravel_multi_index(idx, shape), np.ravel_multi_index(tuple(v // 2 for v in idx), cs)] = 1.0
    Ac = O.face_sparse(O.galerkin_coarsen(op)).toarray()
    np.testing.assert_allclose(Ac, P.T @ A @ P / 8.0, rtol=1e-12, atol=1e-12)
    # R = P^T/8 and prolongation = P
    r = np.random.default_rng(7).standard_normal(shape)
    np.testing.assert_allclose(O.restrict_sum8(r).ravel(), P.T @ r.ravel() / 8.0, atol=1e-14)
    e = np.random.default_rng(8).standard_normal(cs)
    np.testing.assert_allclose(O.prolong_agg(e, shape).ravel(), P @ e.ravel(), atol=1e-15)


def test_level_shapes_rule():
    assert O.level_shapes((64, 64, 64)) == [(64,) * 3, (32,) * 3, (16,) * 3, (8,) * 3, (4,) * 3]
    assert O.level_shapes((48,) * 3)[-1] == (3, 3, 3)
    assert O.level_shapes((208,) * 3)[-2:] == [(7,) * 3, (4,) * 3]
    assert O.level_shapes((64, 256, 256))[-1] == (1, 4, 4)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_fcg_solve_matches_fp64_direct(storage):
    op, b = O.make_anisotropic_problem((16, 16, 16), 11, 1.5, 2)
    res = O.fcg_solve(op, b, tol=1e-8, storage=storage)
    assert res.status == "converged"
    xd = O.solve_fp64_direct(op.astype("f32"), b.astype(np.float32))
    err = np.linalg.norm(res.x - xd) / np.linalg.norm(xd)
    assert err < 1e-6
    # the relative residual recorded is the true residual of x
    r = b.astype(np.float32).astype(np.float64) - O.face_apply(op.astype("f32"), res.x)
    assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(res.rel_history[-1], rel=1e-3)


def test_vcycle_wider_output_changes_only_the_last_rounding():
    # z kept in FP32 over a BF16 cycle: rounding it to BF16 gives the BF16 cycle
    # bit for bit (BF16 rounding goes through FP32 in both), and it is not BF16
    op, b = O.make_anisotropic_problem((16, 12, 20), 5, 1.0, 2)
    h = O.build_hierarchy(op, "bf16")
    for nu2 in (1, 2, 3):
        zb = O.vcycle(h, O.round_storage(b, "bf16"), nu2=nu2, alpha=1.6)
        zf = O.vcycle(h, O.round_storage(b, "bf16"), nu2=nu2, alpha=1.6, out_storage="f32")
        np.testing.assert_array_equal(O.round_storage(zf, "bf16"), zb)
        assert np.any(zf != zb)
        np.testing.assert_array_equal(O.round_storage(zf, "f32"), zf)


def test_fcg_wider_direction_converges_to_fp64_direct():
    op, b = O.make_anisotropic_problem((16, 16, 16), 11, 1.5, 2)
    res = O.fcg_solve(op, b, tol=1e-8, storage="bf16", outer="f32")
    ref = O.fcg_solve(op, b, tol=1e-8, storage="bf16")
    assert res.status == "converged" and res.iters <= ref.iters
    xd = O.solve_fp64_direct(op.astype("f32"), b.astype(np.float32))
    assert np.linalg.norm(res.x - xd) / np.linalg.norm(xd) < 1e-6


def test_vcycle_is_symmetric_positive_definite():
    # the FP64 V(2,2) cycle with exact coarse solve is an SPD preconditioner
    op, _ = _rand_faces((6, 6, 6), seed=9)
    h = O.build_hierarchy(op, "f64")
    n = int(np.prod(op.shape))
    M = np.zeros((n, n))
    for q in range(n):
        e = np.zeros(n); e[q] = 1.0
        M[:, q] = O.vcycle(h, e.reshape(op.shape), alpha=1.6).ravel()
    assert np.abs(M - M.T).max() < 1e-10 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_coarsest_level_is_solved_exactly():
    op, _ = _rand_faces((8, 8, 8), seed=10)
    h = O.build_hierarchy(op, "f64")
    assert h.ops[-1].shape == (4, 4, 4)
    b = np.random.default_rng(11).standard_normal((4, 4, 4))
    e = O.vcycle(h, b, level=len(h.ops) - 1)
    np.testing.assert_allclose(O.face_apply(h.ops[-1], e), b, atol=1e-12)


def test_27_point_closed_forms():
    n = 7
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inner = (slice(1, -1),) * 3
    dif = O.diffusion27_kernel(dt=1.0, Re=1.0, dx=1.0)
    assert abs(dif.sum()) < 1e-15
    for f in (i, j, k):
        np.testing.assert_allclose(O.conv_same((f ** 2).astype(float), dif)[inner], 2.0, atol=1e-12)
    assert O.poisson27_kernel()[1, 1, 1] == pytest.approx(88 / 26)
    # smoothed central differences of linear fields: x-kernel -> +df/dx, y/z kernels -> -df/dy, -df/dz
    kx, ky, kz = O.central_x_kernel(), O.central_y_kernel(), O.central_z_kernel()
    np.testing.assert_allclose(O.conv_same(i.astype(float), kx)[inner], 1.0, atol=1e-12)
    np.testing.assert_allclose(O.conv_same(j.astype(float), ky)[inner], -1.0, atol=1e-12)
    np.testing.assert_allclose(O.conv_same(k.astype(float), kz)[inner], -1.0, atol=1e-12)
    ax, _, _ = O.advection_kernels(dt=0.4, dx=1.0)
    np.testing.assert_allclose(O.conv_same(i.astype(float), ax)[inner], 0.4 * 0.5, atol=1e-12)


def test_boundary_functions_closed_forms():
    rng = np.random.default_rng(12)
    u, v, w = (rng.standard_normal((4, 5, 6)) for _ in range(3))
    uo, vo, wo = O.boundary_condition_velocity(u, v, w, ub=1.5)
    assert np.all(uo[1:-1, 1:-1, 0] == 1.5) and np.all(vo[1:-1, :, -1] == 0)
    np.testing.assert_array_equal(uo[1:-1, 0, :], u[1:-1, 1, :])     # copies read the input
    np.testing.assert_array_equal(uo[-1], u[-2])
    np.testing.assert_array_equal(vo[-1], v[-2])
    assert np.all(wo[-1] == 0) and np.all(uo[0] == 0)
    p = O.boundary_condition_pressure(u)
    assert np.all(p[1:-1, 1:-1, -1] == 0)
    np.testing.assert_array_equal(p[1:-1, 1:-1, 0], u[1:-1, 1:-1, 1])
    np.testing.assert_array_equal(p[0], u[1])
    b = O.boundary_condition_source(u)
    np.testing.assert_array_equal(b[:, :, 0], u[:, :, 1])
    a = O.boundary_condition_indicator(u)
    assert a[0].any() == 0 and a[1:-1, 1:-1, 1:-1].all()
    (f,) = O.bluff_body((u,), np.full(u.shape, 1e5), 0.4)
    np.testing.assert_allclose(f, u / (1 + 0.4e5))
    t = O.boundary_condition_top(u)
    np.testing.assert_array_equal(t[0], u[1])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = O.round_bf16(x)
    np.testing.assert_array_equal(r[:5], [1.0, 1.0, 1.0 + 2 ** -7 * 2, 1.0, -2.5])


@pytest.mark.parametrize("shape,seed,sig", [((12, 10, 8), 1, 1.0), ((16, 16, 16), 2, 2.0),
                                            ((9, 20, 14), 3, 0.5)])
def test_fp64_pcg_matches_direct_solve(shape, seed, sig):
    """The FP64 reference of the large-instance fixtures (solve_fp64_pcg) against
    scipy's sparse direct solve."""
    sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, 2)
    op = O.harmonic_faces(sx, sy, sz).astype("f32")
    x = O.solve_fp64_pcg(op, b, tol=1e-13)
    xd = O.solve_fp64_direct(op, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    r = b - O.face_apply(op, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-12


def test_fp64_reference_fixtures_are_consistent():
    """The committed FP64 reference samples: each solve's recorded residual is at
    the FP64 level and its sampled cell indices are distinct and inside the grid
    it names."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "fp64_reference_solves.npz"))
    for q in range(2):
        assert gold[f"case{q}_relres"][0] < 1e-12
        D, H, W, seed, sig, nd = gold[f"case{q}_spec"]
        assert gold[f"case{q}_idx"].max() < D * H * W
        assert np.all(np.diff(gold[f"case{q}_idx"]) > 0)


def test_solve_refined_is_the_stationary_iteration():
    """Defect correction with one V-cycle per step (solve_refined, inner_cycles=1) in
    FP64 is the stationary iteration e <- (I - M A) e of the V-cycle's matrix M
    (assembled column by column from unit right-hand sides): its residual history
    is ||A (I - M A)^k x*|| / ||b|| with x* the sparse direct solution."""
    sx, sy, sz, b = O.make_anisotropic_fields((8, 6, 4), 4, 1.0, 1)
    op = O.harmonic_faces(sx, sy, sz).astype("f32")   # the outer residual's FP32 faces
    b = b.astype(np.float64)
    h = O.build_hierarchy(op, "f64")
    h.ops = [op] + h.ops[1:]
    n = b.size
    M = np.zeros((n, n))
    for q in range(n):
        e = np.zeros(n); e[q] = 1.0
        M[:, q] = O.vcycle(h, e.reshape(b.shape), alpha=1.6).ravel()
    A = O.face_sparse(op).toarray()
    xs = np.linalg.solve(A, b.ravel())
    G = np.eye(n) - M @ A
    pred, err = [], xs.copy()
    for _ in range(12):
        pred.append(np.linalg.norm(A @ err) / np.linalg.norm(b))
        err = G @ err
    res = O.solve_refined(op, b, tol=1e-30, storage="f64", inner_cycles=1, alpha=1.6,
                          max_outer=11, hierarchy=h)
    np.testing.assert_allclose(res.rel_history[:12], pred, rtol=1e-9, atol=1e-15)
    # and it converges to the direct solution at the contraction rate of G
    rho = max(abs(np.linalg.eigvals(G)))
    assert rho < 0.5
    res = O.solve_refined(op, b, tol=1e-11, storage="f64", inner_cycles=1, alpha=1.6,
                          max_outer=100, hierarchy=h)
    assert res.converged
    assert np.linalg.norm(res.x.ravel() - xs) / np.linalg.norm(xs) < 1e-9


def test_fcg_is_invariant_to_the_residual_scale():
    """The FP16 path's per-iteration power-of-2 scaling of the V-cycle's
    right-hand side, with r.z divided by the scale, is the same iteration: in
    FP64 storage the scaled and unscaled solves take the same iterations and
    agree to rounding (dividing r.z by the square of the scale, or not at all,
    changes beta and the iterates)."""
    sx, sy, sz, b = O.make_anisotropic_fields((12, 10, 8), 3, 1.5, 2)
    op = O.harmonic_faces(sx, sy, sz)
    a = O.fcg_solve(op, b, tol=1e-10, storage="f64", scaled=False)
    for e in (-12, -4, 6):
        c = O.fcg_solve(op, b, tol=1e-10, storage="f64", scaled=True, f16_rms_exp=e)
        assert c.iters == a.iters and c.status == "converged"
        assert np.linalg.norm(c.x - a.x) / np.linalg.norm(a.x) < 1e-10
        np.testing.assert_allclose(c.rel_history, a.rel_history, rtol=1e-7, atol=1e-14)


def test_fcg_f16_storage_reaches_the_fp64_solution():
    sx, sy, sz, b = O.make_anisotropic_fields((16, 12, 16), 4, 1.0, 1)
    op = O.harmonic_faces(sx, sy, sz)
    res = O.fcg_solve(op, b, tol=1e-8, storage="f16")
    assert res.status == "converged"
    xd = O.solve_fp64_direct(op.astype("f32"), b.astype(np.float32))
    assert np.linalg.norm(res.x - xd) / np.linalg.norm(xd) < 1e-6
    bf = O.fcg_solve(op, b, tol=1e-8, storage="bf16")
    assert res.iters <= bf.iters       # more mantissa than BF16
