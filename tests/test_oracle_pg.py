"""Pins of the Petrov-Galerkin dissipation restatement (CPU): closed forms."""
import numpy as np
import pytest

from oracle import hfm_oracle as O

DT = 0.5


def _k():
    dif = O.diffusion27_kernel(DT)
    return (dif, *O.advection_kernels(DT))


def _interior(a, m=1):
    return a[m:-m, m:-m, m:-m]


def test_uniform_and_linear_flows_are_not_dissipated():
    sh = (6, 7, 8)
    sig = np.zeros(sh, np.float32)
    k, j, i = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in sh], indexing="ij")
    for u in (np.full(sh, 1.5, np.float32), (0.3 + 0.05 * i).astype(np.float32),
              (1.0 - 0.02 * k + 0.03 * j).astype(np.float32)):
        _, ks = O.pg_coefficients(u, u, u, sig, DT, 1e-3, *_k())
        for kk in ks:          # FP32 rounding of dif(f) over eps_k: ~1e-5 of the cap 0.5
            assert np.abs(_interior(kk)).max() < 1e-5
        K = O.pg_dissipation(u, u, u, sig, DT, 1e-3, *_k())
        for kk in K:           # the flux takes dif of k: two cells from the walls
            assert np.abs(_interior(kk, 2)).max() < 1e-6


def test_constant_diffusivity_reduces_to_three_k_dif():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((6, 6, 6)).astype(np.float32)
    dif = O.diffusion27_kernel(DT)
    c = np.float32(0.37)
    got = O.pg_flux(f, np.full(f.shape, c, np.float32), dif)
    want = 3 * c * O.conv_same(f.astype(np.float64), dif)
    np.testing.assert_allclose(_interior(got), _interior(want), rtol=1e-5, atol=1e-6)


def test_odd_in_the_velocity_and_capped():
    rng = np.random.default_rng(2)
    sh = (5, 6, 7)
    u, v, w = (rng.standard_normal(sh).astype(np.float32) for _ in range(3))
    sig = np.zeros(sh, np.float32)
    sig[1:3, 2:4, 2:5] = 1e5
    K = O.pg_dissipation(u, v, w, sig, DT, 1e-3, *_k())
    Kn = O.pg_dissipation(-u, -v, -w, sig, DT, 1e-3, *_k())
    for a, b in zip(K, Kn):
        np.testing.assert_array_equal(a, -b)
    _, ks = O.pg_coefficients(u, v, w, sig, DT, 1e-3, *_k())
    cap = np.float32(0.5) / (np.float32(1.0) + np.float32(DT) * sig)
    for kk in ks:
        assert (kk >= 0).all() and (kk <= cap).all()
        assert (kk == cap).any()          # small eps_k: the cap is reached


def test_known_answer_at_a_point_source():
    # u = v = w = delta at the centre, sigma = 0: at the centre the mean speed
    # (|u|+|v|+|w|)/3 is 1, the central advection stencils have zero centre
    # weight, and dif(delta) is the diffusion kernel's centre weight, so
    # k = 8 * 0.25 * |dif_c| / eps_k (below the cap for a large eps_k).
    sh = (5, 5, 5)
    f = np.zeros(sh, np.float32)
    f[2, 2, 2] = 1
    eps = 1e3
    _, ks = O.pg_coefficients(f, f, f, np.zeros(sh, np.float32), DT, eps, *_k())
    dif_c = -88 / 26 * DT * (1 / 20)
    for kk in ks:
        assert kk[2, 2, 2] == pytest.approx(2 * abs(dif_c) / eps, rel=1e-6)
