"""Pins of the 2D periodic flow oracle (CPU): closed forms, FFT and brute force."""
import math

import numpy as np
import pytest

from oracle import hfm_flow2d as F


def _fft_solve(b):
    """Mean-zero solution of the periodic 5-point A x = b by FFT (A's symbol is
    4 - 2 cos(2 pi k/N) - 2 cos(2 pi l/N))."""
    N = b.shape[0]
    k = 2 * np.pi * np.fft.fftfreq(N)
    sym = 4 - 2 * np.cos(k)[None, :] - 2 * np.cos(k)[:, None]
    bh = np.fft.fft2(b)
    sym[0, 0] = 1.0
    xh = bh / sym
    xh[0, 0] = 0.0
    return np.real(np.fft.ifft2(xh))


def test_apply_periodic_fourier_symbol():
    N = 16
    j, i = np.mgrid[0:N, 0:N]
    for (kx, ky) in ((1, 0), (3, 5), (8, 8)):
        e = np.cos(2 * np.pi * (kx * i + ky * j) / N)
        lam = 4 - 2 * math.cos(2 * math.pi * kx / N) - 2 * math.cos(2 * math.pi * ky / N)
        np.testing.assert_allclose(F.apply_periodic(e), lam * e, atol=1e-12)
    assert np.abs(F.apply_periodic(np.ones((N, N)))).max() == 0.0


def test_coarsest_pinv_is_the_pseudo_inverse():
    for n, s in ((8, 2.0 ** -5), (4, 1.0), (2, 0.5)):
        P = F.coarsest_pinv(n, s)
        m = n * n
        M = np.array([F.apply_periodic(e.reshape(n, n), s).ravel() for e in np.eye(m)]).T
        np.testing.assert_allclose(M @ P @ M, M, atol=1e-10)
        np.testing.assert_allclose(P, P.T, atol=1e-12)
        np.testing.assert_allclose(P.sum(axis=0), 0.0, atol=1e-10)   # mean-zero results
        b = np.random.default_rng(n).standard_normal(m)
        b -= b.mean()
        np.testing.assert_allclose(M @ (P @ b), b, atol=1e-10)


def test_projection_solve_matches_fft():
    N = 64
    mg = F.build_mg(N)
    rng = np.random.default_rng(3)
    b = rng.standard_normal((N, N))
    b = F._r(b, 0, 0) - b.mean()
    res = F.project_solve(mg, b, np.zeros((N, N)), tol=1e-10, maxit=60)
    assert res.status == "converged"
    x = res.phi - res.phi.mean()
    xf = _fft_solve(b)
    assert np.linalg.norm(x - xf) / np.linalg.norm(xf) < 1e-9
    # the stopping rule is the true FP64 residual
    r = b - F.apply_periodic(res.phi)
    assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(res.rel, rel=1e-6)


def test_vcycle_is_spd_on_mean_zero_vectors():
    N = 16
    mg = F.build_mg(N, coarsest=4)
    m = N * N
    Pm = np.eye(m) - 1.0 / m       # projector onto mean-zero vectors
    M = np.zeros((m, m))
    for q in range(m):
        e = Pm[:, q].reshape(N, N).astype(np.float32)
        M[:, q] = F.vcycle(mg, e).astype(np.float64).ravel()
    M = Pm @ M @ Pm
    assert np.abs(M - M.T).max() < 1e-5 * np.abs(M).max()      # FP32 cycle
    w = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert (w > 1e-6).sum() == m - 1


def test_initial_velocity_is_divergence_free():
    N = 64
    u, v = F.random_modes_ic(N, 5)
    h = 2 * np.pi / N
    d = F.divergence(u, v, h)
    grad = np.abs(F._r(u, 0, 1) - u).max() / h
    assert np.abs(d).max() < 1e-5 * grad
    assert math.sqrt(float(np.mean(u.astype(np.float64) ** 2 + v.astype(np.float64) ** 2))) == \
        pytest.approx(1.0, rel=1e-5)
    u2, _ = F.random_modes_ic(N, 5)
    np.testing.assert_array_equal(u, u2)


def _rhs_loops(u, v, nu_h2, inv_h, fu):
    """Brute-force restatement of the momentum right-hand side with explicit loops."""
    N = u.shape[0]
    Fu = np.zeros((N, N))
    Fv = np.zeros((N, N))
    for j in range(N):
        for i in range(N):
            jp, jm, ip, im = (j + 1) % N, (j - 1) % N, (i + 1) % N, (i - 1) % N
            vb = 0.25 * (v[j, im] + v[j, i] + v[jp, im] + v[jp, i])
            ub = 0.25 * (u[jm, i] + u[jm, ip] + u[j, i] + u[j, ip])
            ax = u[j, i] * ((u[j, i] - u[j, im]) if u[j, i] > 0 else (u[j, ip] - u[j, i]))
            ay = vb * ((u[j, i] - u[jm, i]) if vb > 0 else (u[jp, i] - u[j, i]))
            lap = u[j, ip] + u[j, im] + u[jp, i] + u[jm, i] - 4 * u[j, i]
            Fu[j, i] = nu_h2 * lap - inv_h * (ax + ay) + fu[j, 0]
            bx = ub * ((v[j, i] - v[j, im]) if ub > 0 else (v[j, ip] - v[j, i]))
            by = v[j, i] * ((v[j, i] - v[jm, i]) if v[j, i] > 0 else (v[jp, i] - v[j, i]))
            lap = v[j, ip] + v[j, im] + v[jp, i] + v[jm, i] - 4 * v[j, i]
            Fv[j, i] = nu_h2 * lap - inv_h * (bx + by)
    return Fu, Fv


def test_momentum_rhs_brute_force():
    N = 8
    rng = np.random.default_rng(11)
    u = rng.standard_normal((N, N)).astype(np.float32)
    v = rng.standard_normal((N, N)).astype(np.float32)
    fu = F.forcing_u(N)
    Fu, Fv = F.momentum_rhs(u, v, 0.3, 1.7, fu)
    Gu, Gv = _rhs_loops(u.astype(np.float64), v.astype(np.float64), 0.3, 1.7, fu.astype(np.float64))
    np.testing.assert_allclose(Fu, Gu, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(Fv, Gv, rtol=1e-5, atol=1e-5)


def test_laminar_kolmogorov_state_is_a_fixed_point():
    N, nu, dt = 64, 0.01, 0.005
    mg = F.build_mg(N, coarsest=8)
    u, v = F.laminar_state(N, nu)
    st = F.FlowState(u.copy(), v.copy(), np.zeros((N, N)))
    for _ in range(10):
        F.step(st, mg, dt, nu, F.forcing_u(N))
    assert np.abs(st.u - u).max() < 2e-5 * np.abs(u).max()
    assert np.abs(st.v).max() < 1e-6 * np.abs(u).max()


def test_taylor_green_decays_at_the_discrete_viscous_rate():
    # amplitude small enough that the upwind advection is negligible; the
    # two-stage update of the 5-point diffusion multiplies the mode by
    # 1 - z + z^2/2 per step, z = nu lambda_h dt.
    N, nu, dt, n = 32, 0.1, 0.01, 60
    h = 2 * np.pi / N
    mg = F.build_mg(N, coarsest=8)
    u0, v0 = F.taylor_green(N, 1e-3)
    st = F.FlowState(u0.copy(), v0.copy(), np.zeros((N, N)))
    zero = np.zeros((N, 1), np.float32)
    for _ in range(n):
        F.step(st, mg, dt, nu, zero)
    lam = 2 * (2 - 2 * math.cos(h)) / h ** 2
    z = nu * lam * dt
    expect = (1 - z + z * z / 2) ** n
    got = float(np.sum(st.u.astype(np.float64) * u0) / np.sum(u0.astype(np.float64) * u0))
    assert math.log(got) / math.log(expect) == pytest.approx(1.0, rel=0.01)
    assert abs(expect - 1) > 0.1           # the decay is visible


def test_projection_removes_a_gradient_and_keeps_a_solenoidal_field():
    N = 32
    h = 2 * np.pi / N
    mg = F.build_mg(N, coarsest=8)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal((N, N))
    us = (psi - F._r(psi, 0, -1)) / h         # pure gradient on the MAC points
    vs = (psi - F._r(psi, -1, 0)) / h
    b = F.divergence_rhs(us.astype(np.float32), vs.astype(np.float32), h)
    res = F.project_solve(mg, b, np.zeros((N, N)), tol=1e-9)
    u = us - (res.phi - F._r(res.phi, 0, -1)) / h
    assert np.abs(u).max() < 1e-5 * np.abs(us).max()
    u1, v1 = F.random_modes_ic(N, 2)
    b = F.divergence_rhs(u1, v1, h)
    res = F.project_solve(mg, b, np.zeros((N, N)), tol=1e-6)
    assert np.abs(res.phi).max() < 1e-4 * h * np.abs(u1).max() * N


def test_substeps_follow_the_cfl_rule():
    N = 64
    h = 2 * np.pi / N
    u = np.full((N, N), 2.0, np.float32)
    v = np.full((N, N), -1.0, np.float32)
    # (2 + 1) * 0.005 / h / 0.5 = 0.306 -> 1 substep; dt 0.02 -> 1.22 -> 2
    assert F.substeps(u, v, 0.005, h) == 1
    assert F.substeps(u, v, 0.02, h) == 2
    assert F.substeps(u, v, 0.1, h) == math.ceil(3 * 0.1 / h / 0.5)
