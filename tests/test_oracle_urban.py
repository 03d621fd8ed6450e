"""The oracle's urban / INHALE time steps and the 2D explicit update against
independent restatements and closed forms (CPU).

The pressure cycle of the urban drivers is checked against a dense matrix
restatement built here from Kronecker products (restriction, prolongation)
and an explicit assembly of the 27-point Poisson matrix, written without the
oracle's layer functions; the form with a fine-level sweep of the correction
must fail that check."""
import numpy as np
import pytest

from oracle import hfm_oracle as O


# ---------------------------------------------------------------- dense restatement

def dense_conv_matrix(shape, K):
    """Matrix of the SAME cross-correlation y[c] = sum_t K[t] x[c + t - 1] (zero ghosts)."""
    D, H, W = shape
    n = D * H * W
    A = np.zeros((n, n))
    for k in range(D):
        for j in range(H):
            for i in range(W):
                row = (k * H + j) * W + i
                for a in range(3):
                    for b in range(3):
                        for c in range(3):
                            kk, jj, ii = k + a - 1, j + b - 1, i + c - 1
                            if 0 <= kk < D and 0 <= jj < H and 0 <= ii < W:
                                A[row, (kk * H + jj) * W + ii] += K[a, b, c]
    return A


def mean_1d(n):
    R = np.zeros((n // 2, n))
    for q in range(n // 2):
        R[q, 2 * q] = R[q, 2 * q + 1] = 0.5
    return R


def repeat_1d(n):
    P = np.zeros((2 * n, n))
    for q in range(2 * n):
        P[q, q // 2] = 1.0
    return P


def kron3(a, b, c):
    return np.kron(np.kron(a, b), c)


def outlet_mask(shape):
    z = np.ones(shape)
    z[..., -1] = 0.0
    return np.diag(z.ravel())


def dense_urban_cycle(x, b, shapes, K):
    """One pressure cycle of UM/AI_HFM_v3.ipynb:2570-2613 as matrices: r = Z (A x - b);
    restrict to every level; w = J(0, r) at the coarsest; w = J(P w, r_l) on the way
    up to level 1; the finest correction is P w with no sweep; x = Z (x - P w); one
    fine sweep."""
    d = K[1, 1, 1]
    A = [dense_conv_matrix(s, K) for s in shapes]
    R = [kron3(*(mean_1d(n) for n in s)) for s in shapes[:-1]]        # level l -> l + 1
    P = [kron3(*(repeat_1d(n) for n in s)) for s in shapes[1:]]        # level l + 1 -> l
    Z = outlet_mask(shapes[0])
    J = lambda lev, w, r: w - A[lev] @ w / d + r / d  # noqa: E731
    xv, bv = x.ravel(), b.ravel()
    rs = [Z @ (A[0] @ xv - bv)]
    for lev in range(len(shapes) - 1):
        rs.append(R[lev] @ rs[-1])
    L = len(shapes) - 1
    w = J(L, np.zeros(rs[L].size), rs[L])
    for lev in range(L - 1, 0, -1):
        w = J(lev, P[lev] @ w, rs[lev])
    w0 = P[0] @ w
    x1 = Z @ (xv - w0)
    x2 = J(0, x1, bv)
    return x2.reshape(x.shape), w0.reshape(x.shape)


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


@pytest.mark.parametrize("shape", [(4, 4, 4), (2, 4, 8), (4, 8, 4)])
def test_urban_cycle_matches_dense_restatement(shape):
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape)
    b = rng.standard_normal(shape)
    K = O.poisson27_kernel()
    shapes = O.urban_levels(shape)
    ref, w0 = dense_urban_cycle(x, b, shapes, K)
    ops = [lambda v: O.conv_same(v, K)] * len(shapes)
    rec = {}
    got = O.sawtooth_cycle(x, b, ops, [K[1, 1, 1]] * len(shapes), zero_outlet=True,
                           fine_sweep=False, record=rec)
    assert rel(got, ref) < 1e-12
    assert rel(rec[0], w0) < 1e-12          # the guard's w_512: the prolonged correction
    # the form that also sweeps the fine correction is not the reference's cycle
    other = O.sawtooth_cycle(x, b, ops, [K[1, 1, 1]] * len(shapes), zero_outlet=True)
    assert rel(other, ref) > 1e-3


def test_urban_pressure_cycles_fp32_match_dense():
    rng = np.random.default_rng(12)
    shape = (4, 8, 8)
    p = rng.standard_normal(shape).astype(np.float32)
    b = rng.standard_normal(shape).astype(np.float32)
    K = O.poisson27_kernel()
    shapes = O.urban_levels(shape)
    ref = p.astype(np.float64)
    for _ in range(3):
        ref, w0 = dense_urban_cycle(ref, b.astype(np.float64), shapes, K)
    rec = {}
    got = O.urban_pressure_cycles(p, b, multi_itr=3, record=rec)
    assert got.dtype == np.float32
    assert rel(got, ref) < 1e-5
    assert rel(rec[0], w0) < 1e-5


def test_script_twin_coarsest_level_is_two():
    """ARTs/naiver_stokes_3D.py:648: the hierarchy stops at 2^3 and the cycle starts
    from w_2 = 0 there."""
    assert O.urban_levels((16, 16, 16), coarsest=2)[-1] == (2, 2, 2)
    assert O.urban_levels((16, 16, 16))[-1] == (1, 1, 1)
    rng = np.random.default_rng(13)
    shape = (8, 8, 8)
    x, b = rng.standard_normal(shape), rng.standard_normal(shape)
    K = O.poisson27_kernel()
    shapes = O.urban_levels(shape, coarsest=2)
    ref, _ = dense_urban_cycle(x, b, shapes, K)
    got = O.sawtooth_cycle(x, b, [lambda v: O.conv_same(v, K)] * len(shapes),
                           [K[1, 1, 1]] * len(shapes), zero_outlet=True, fine_sweep=False)
    assert rel(got, ref) < 1e-12


def test_j_itr_sweeps_every_level_below_the_finest():
    """j_itr = 2: two sweeps after each prolongation (and at the coarsest), still none
    on the finest correction."""
    rng = np.random.default_rng(14)
    shape = (4, 4, 4)
    x, b = rng.standard_normal(shape), rng.standard_normal(shape)
    K = O.poisson27_kernel()
    d = K[1, 1, 1]
    shapes = O.urban_levels(shape)
    A = [dense_conv_matrix(s, K) for s in shapes]
    R = [kron3(*(mean_1d(n) for n in s)) for s in shapes[:-1]]
    P = [kron3(*(repeat_1d(n) for n in s)) for s in shapes[1:]]
    Z = outlet_mask(shape)
    J = lambda lev, w, r: w - A[lev] @ w / d + r / d  # noqa: E731
    rs = [Z @ (A[0] @ x.ravel() - b.ravel())]
    for lev in range(len(shapes) - 1):
        rs.append(R[lev] @ rs[-1])
    w = J(2, J(2, np.zeros(1), rs[2]), rs[2])
    w = P[1] @ w
    w = J(1, J(1, w, rs[1]), rs[1])
    x1 = Z @ (x.ravel() - P[0] @ w)
    ref = J(0, x1, b.ravel()).reshape(shape)
    got = O.sawtooth_cycle(x, b, [lambda v: O.conv_same(v, K)] * 3, [d] * 3, zero_outlet=True,
                           fine_sweep=False, j_itr=2)
    assert rel(got, ref) < 1e-12


# ---------------------------------------------------------------- stage invariants

def _state(shape, seed, ub=1.0):
    rng = np.random.default_rng(seed)
    u = (ub + 0.1 * rng.standard_normal(shape)).astype(np.float32)
    v, w, p, a = (0.1 * rng.standard_normal(shape).astype(np.float32) for _ in range(4))
    sig = np.zeros(shape, np.float32)
    sig[1:3, 2:4, 3:5] = 1e5
    return u, v, w, p, a, sig


@pytest.mark.parametrize("step", [O.urban_step, O.inhale_step])
def test_zero_state_stays_zero(step):
    shape = (4, 8, 8)
    z = np.zeros(shape, np.float32)
    sig = _state(shape, 0)[-1]
    out = step(z, z, z, z, z, sig, ub=0.0, multi_itr=2)
    for f in out:
        assert np.all(np.asarray(f) == 0)


def test_bluff_body_sigma_zero_is_identity():
    rng = np.random.default_rng(3)
    f = rng.standard_normal((3, 4, 5)).astype(np.float32)
    for g in O.bluff_body((f, 2 * f), np.zeros_like(f), 0.4):
        assert g.dtype == np.float32
    out = O.bluff_body((f,), np.zeros_like(f), 0.4)[0]
    np.testing.assert_array_equal(out, f)
    big = O.bluff_body((f,), np.full_like(f, 1e5), np.float32(0.4))[0]
    np.testing.assert_allclose(big, f / np.float32(1 + 0.4 * 1e5), rtol=1e-7)


def _mirror_y(state):
    u, v, w, p, a = state[:5]
    m = lambda t: np.ascontiguousarray(t[:, ::-1, :])  # noqa: E731
    return m(u), -m(v), m(w), m(p), m(a)


@pytest.mark.parametrize("step", [O.urban_step, O.inhale_step])
def test_step_is_mirror_equivariant_in_y(step):
    """The y kernels differentiate toward decreasing j and the y faces take the same
    fill on both sides, so reflecting j (and negating v) commutes with the step: a
    kernel applied along the wrong axis, a one-sided y boundary or a transposed
    operand breaks this."""
    shape = (4, 8, 8)
    u, v, w, p, a, sig = _state(shape, 5)
    sig_m = np.ascontiguousarray(sig[:, ::-1, :])
    out = step(u, v, w, p, a, sig, multi_itr=2, lscalar=True)
    out_m = step(*_mirror_y((u, v, w, p, a)), sig_m, multi_itr=2, lscalar=True)
    for g, h in zip(_mirror_y(out), out_m):
        assert rel(h, g) < 1e-5, rel(h, g)


@pytest.mark.parametrize("step", [O.urban_step, O.inhale_step])
def test_step_guard_reads_prolonged_correction(step):
    shape = (4, 8, 8)
    u, v, w, p, a, sig = _state(shape, 6)
    out = step(u, v, w, p, a, sig, multi_itr=1)
    w_fine = out[5]
    assert w_fine.shape == shape
    # nearest prolongation: every 2x2x2 block of w_512 holds one value
    blocks = w_fine.reshape(2, 2, 4, 2, 4, 2)
    assert np.all(blocks == blocks[:, :1, :, :1, :, :1])


# ---------------------------------------------------------------- 2D explicit update

@pytest.mark.parametrize("name", ["AD2D_UPWIND", "AD2D_CENTRAL"])
def test_ad2d_kernels_sum_to_zero(name):
    K = getattr(O, name)
    assert abs(K.sum()) < 1e-15
    x = np.full((9, 9), 3.0)
    np.testing.assert_allclose(O.ad2d_step(x, K)[2:-2, 2:-2], x[2:-2, 2:-2], atol=1e-13)


@pytest.mark.parametrize("name", ["AD2D_UPWIND", "AD2D_CENTRAL"])
def test_ad2d_step_moments(name):
    """For S = I + K + K^2/2 with sum(K) = 0: mass is kept, the centre of mass moves by
    the first moment m1 of K per step, and the variance along an axis grows by
    m2 - m1^2 + m1^2 = m2 (the second moment of K), away from the boundary."""
    K = getattr(O, name)
    off = np.array([-1.0, 0.0, 1.0])
    # y[j, i] = sum K[a, c] x[j + a - 1, i + c - 1]: mass at p moves to p - (offset)
    m1 = np.array([-(K.sum(axis=1) * off).sum(), -(K.sum(axis=0) * off).sum()])
    m2 = np.array([(K.sum(axis=1) * off ** 2).sum(), (K.sum(axis=0) * off ** 2).sum()])
    x = np.zeros((41, 41))
    x[20, 20] = 1.0
    nsteps = 5
    y = x
    for _ in range(nsteps):
        y = O.ad2d_step(y, K)
    jj, ii = np.meshgrid(np.arange(41.0), np.arange(41.0), indexing="ij")
    mass = y.sum()
    assert abs(mass - 1.0) < 1e-13
    cj, ci = (y * jj).sum(), (y * ii).sum()
    np.testing.assert_allclose([cj - 20, ci - 20], nsteps * m1, atol=1e-12)
    vj = (y * (jj - cj) ** 2).sum()
    vi = (y * (ii - ci) ** 2).sum()
    np.testing.assert_allclose([vj, vi], nsteps * m2, atol=1e-11)
    # the upwind kernel carries the fixed velocity u dt/dx = 0.1 along +x and rows
    if name == "AD2D_UPWIND":
        np.testing.assert_allclose(m1, [-0.1, 0.1], atol=1e-15)


# ---------------------------------------------------------------- Multigrid_cnn 2D cycle

def test_prolong_upsum_is_upsampling_then_same_conv():
    """prolong_upsum = torch nearest upsampling, one zero of padding after the end of
    each axis, and a 2x2 all-ones cross-correlation (TensorFlow's SAME padding of an
    even window); a constant coarse field gives 4 c inside, 2 c on the last row or
    column and c in the last corner."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(4)
    c = rng.standard_normal((5, 7))
    t = torch.from_numpy(c)[None, None]
    up = F.interpolate(t, scale_factor=2, mode="nearest")
    ref = F.conv2d(F.pad(up, (0, 1, 0, 1)), torch.ones(1, 1, 2, 2, dtype=torch.float64))[0, 0]
    np.testing.assert_allclose(O.prolong_upsum(c), ref.numpy(), rtol=1e-14, atol=1e-14)
    y = O.prolong_upsum(np.full((3, 3), 2.0))
    assert np.all(y[:-1, :-1] == 8.0) and np.all(y[-1, :-1] == 4.0) and np.all(y[:-1, -1] == 4.0)
    assert y[-1, -1] == 2.0


def _upsum_1d(n):
    """(2n x n): nearest repeat then the window {i, i+1} sum, zero past the end."""
    S = np.zeros((2 * n, 2 * n))
    for i in range(2 * n):
        S[i, i] = 1.0
        if i + 1 < 2 * n:
            S[i, i + 1] = 1.0
    return S @ repeat_1d(n)


def dense_conv2d_matrix(shape, K):
    H, W = shape
    A = np.zeros((H * W, H * W))
    for j in range(H):
        for i in range(W):
            for a in range(3):
                for b in range(3):
                    jj, ii = j + a - 1, i + b - 1
                    if 0 <= jj < H and 0 <= ii < W:
                        A[j * W + i, jj * W + ii] += K[a, b]
    return A


@pytest.mark.parametrize("shape", [(4, 4), (8, 4)])
def test_mgcnn_cycle_matches_dense_restatement(shape):
    """ARTn/Multigrid_cnn.ipynb cell 21 as matrices: pre-sweep, r = A x - b, 2x2 mean
    restriction to every level, j_itr = 2 sweeps from zero at the coarsest and after
    each prolongation below the finest, one sweep at the finest, x -= w, post-sweep;
    the smoother w - c A w + c r with c = 1/107."""
    rng = np.random.default_rng(15)
    x, b = rng.standard_normal(shape), rng.standard_normal(shape)
    K, c = O.MGCNN_A, O.MGCNN_BIAS
    shapes = [shape]
    while min(shapes[-1]) > 1:
        shapes.append(tuple(n // 2 for n in shapes[-1]))
    A = [dense_conv2d_matrix(s, K) for s in shapes]
    R = [np.kron(mean_1d(s[0]), mean_1d(s[1])) for s in shapes[:-1]]
    P = [np.kron(_upsum_1d(s[0]), _upsum_1d(s[1])) for s in shapes[1:]]
    J = lambda lev, w, r: w - c * (A[lev] @ w) + c * r  # noqa: E731
    xv, bv = x.ravel(), b.ravel()
    xv = J(0, xv, bv)
    rs = [A[0] @ xv - bv]
    for lev in range(len(shapes) - 1):
        rs.append(R[lev] @ rs[-1])
    L = len(shapes) - 1
    w = np.zeros(rs[L].size)
    for _ in range(2):
        w = J(L, w, rs[L])
    for lev in range(L - 1, 0, -1):
        w = P[lev] @ w
        for _ in range(2):
            w = J(lev, w, rs[lev])
    w = J(0, P[0] @ w, rs[0])
    ref = J(0, xv - w, bv).reshape(shape)
    got, norm = O.mgcnn_cycle(x, b)
    assert rel(got, ref) < 1e-12
    assert abs(norm - rs[0].sum() / rs[0].size) < 1e-12
    # padding before the start instead of after the end is another prolongation
    orig = O.prolong_upsum
    try:
        O.prolong_upsum = lambda w: np.flip(orig(np.flip(w)))
        other, _ = O.mgcnn_cycle(x, b)
    finally:
        O.prolong_upsum = orig
    assert rel(other, ref) > 1e-3


def test_mgcnn_initial_condition():
    """Cells 14-16: the interior Gaussian times 0.8 peaks near the centre, the 0.2
    lattice covers 10 of every 16 cells."""
    v = O.mgcnn_initial_condition(128)
    blob = v[0]                          # row 0: the Gaussian is zero on the edge rows
    assert np.isclose(blob[::4], 0.2).all() and np.isclose(blob[1::4], 0.0).all()
    assert abs((v > 0.19).mean() - 10 / 16) < 0.05
    assert 0.95 < v.max() < 1.0 + 1e-12
