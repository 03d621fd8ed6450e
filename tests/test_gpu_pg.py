"""Petrov-Galerkin dissipation kernels and the INHALE demo step against the oracle (GPU)."""
import numpy as np
import pytest

from oracle import hfm_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def T(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)


def H(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _flow_fields(sh, seed):
    rng = np.random.default_rng(seed)
    u = (1.0 + 0.3 * rng.standard_normal(sh)).astype(np.float32)
    v, w = (0.2 * rng.standard_normal(sh).astype(np.float32) for _ in range(2))
    sig = np.zeros(sh, np.float32)
    sig[1:4, 3:6, 5:9] = 1e5
    return u, v, w, sig


@pytest.mark.parametrize("eps", [1e-3, 10.0])
def test_pg_dissipation_matches_oracle(dev, eps):
    from hfm import layers
    from hfm.timestep import urban_kernels
    dt = 0.5
    k = urban_kernels(dt, 1.0, 1 / 20)
    K = (k["dif"], *k["adv"])
    sh = (6, 10, 24)
    u, v, w, sig = _flow_fields(sh, 4)
    got = layers.petrov_galerkin_dissipation(T(u, dev), T(v, dev), T(w, dev), eps, T(sig, dev),
                                             dt, K)
    ref = O.pg_dissipation(u, v, w, sig, dt, eps, *K)
    for g, r in zip(got, ref):
        assert rel(H(g), r) < 1e-5


def test_dissipative_advection_matches_oracle(dev):
    from hfm import layers
    from hfm.timestep import urban_kernels
    k = urban_kernels(0.5, 1.0, 1 / 20)
    ax, ay, az = k["adv"]
    sh = (5, 12, 20)
    rng = np.random.default_rng(5)
    u, v, w, b0, b1, b2, d0, d1, d2 = (rng.standard_normal(sh).astype(np.float32) for _ in range(9))
    outs = layers.advect_dissipative([T(x, dev) for x in (u, v, w)], [T(x, dev) for x in (u, v, w)],
                                     [T(x, dev) for x in (b0, b1, b2)],
                                     [T(x, dev) for x in (d0, d1, d2)], 0.5, (ax, ay, az))
    C = lambda x, Kk: O.conv_same(x.astype(np.float64), Kk)  # noqa: E731
    for g, phi, b, d in zip(outs, (u, v, w), (b0, b1, b2), (d0, d1, d2)):
        ref = 0.5 * (((d - u * C(phi, ax)) - v * C(phi, ay)) - w * C(phi, az)) + b
        assert rel(H(g), ref) < 1e-6


@pytest.mark.parametrize("sh", [(8, 16, 16), (4, 16, 16)])
def test_inhale_steps_match_oracle(dev, sh):
    # (4, 16, 16): a flat box as in the demo (nz = nx / ratio; coarsest 1 x 4 x 4)
    # UM/INHALE_demo.ipynb cell 24: the prolonged fine correction w_256 applied
    # with no fine-level sweep (:13454-13458), the guard max|w_256| > nsafe
    from hfm.timestep import InhaleFlow
    rng = np.random.default_rng(2)
    u = (1.0 + 0.05 * rng.standard_normal(sh)).astype(np.float32)
    v, w, p = (0.05 * rng.standard_normal(sh).astype(np.float32) for _ in range(3))
    a = np.zeros(sh, np.float32)
    sig = np.zeros(sh, np.float32); sig[1:3, 6:9, 6:9] = 1e5
    flow = InhaleFlow(sh, sigma=T(sig, dev), multi_itr=3)
    g = [T(x, dev) for x in (u, v, w, p, a)]
    o = (u, v, w, p, a)
    for _ in range(3):
        g = flow.step(*g)
        res = O.inhale_step(*o, sig, multi_itr=3)
        o = res[:5]
    for gi, oi in zip(g[:4], o[:4]):
        assert rel(H(gi), oi) < 1e-5
    assert rel(H(flow.last_w), res[5]) < 1e-5
    assert not flow.diverged()
