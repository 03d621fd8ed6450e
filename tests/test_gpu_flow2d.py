"""Parity of the 2D periodic flow kernel with the oracle (GPU, through the C ABI)."""
import numpy as np
import pytest

from oracle import hfm_flow2d as F

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _flow(dev, N, nflows=1, **kw):
    from hfm.flow2d import KolmogorovFlow
    nlev = kw.pop("nlev", int(np.log2(N // 8)) + 1)      # down to 8 x 8
    return KolmogorovFlow(N=N, nflows=nflows, nlev=nlev, device=dev, **kw)


def _oracle_run(N, seed, steps, nu=0.01, dt=0.005, tol=1e-6):
    mg = F.build_mg(N)
    u, v = F.random_modes_ic(N, seed)
    st = F.FlowState(u, v, np.zeros((N, N)))
    nsubs = []
    for _ in range(steps):
        n, res = F.step(st, mg, dt, nu, F.forcing_u(N), tol)
        nsubs.append(n)
        assert res.status == "converged"
    return st, nsubs


@pytest.mark.parametrize("N,steps", [(64, 6), (256, 3)])
def test_steps_match_oracle(dev, N, steps):
    import torch
    seeds = (1, 2)
    fl = _flow(dev, N, nflows=len(seeds))
    for q, s in enumerate(seeds):
        u, v = F.random_modes_ic(N, s)
        fl.set_state(q, torch.from_numpy(u), torch.from_numpy(v))
    fl.step(steps)
    st = fl.stats()
    for q, s in enumerate(seeds):
        ref, nsubs = _oracle_run(N, s, steps)
        u, v, phi = (t.cpu().numpy() for t in fl.state(q))
        assert int(st["steps"][q]) == steps
        assert int(st["status"][q]) == 1
        assert int(st["substeps"][q]) == sum(nsubs)
        assert int(st["iters_sum"][q]) == sum(ref.iters)
        assert rel(u, ref.u) < 1e-5 and rel(v, ref.v) < 1e-5
        assert rel(phi - phi.mean(), ref.phi - ref.phi.mean()) < 1e-5
        d = F.divergence(u, v, 2 * np.pi / N)
        g = np.abs(np.roll(u, -1, 1) - u).max() * N / (2 * np.pi)
        assert np.abs(d).max() < 1e-4 * g


def test_projection_only_matches_oracle(dev):
    import torch
    N = 128
    rng = np.random.default_rng(7)
    u = rng.standard_normal((N, N)).astype(np.float32)
    v = rng.standard_normal((N, N)).astype(np.float32)
    fl = _flow(dev, N)
    fl.set_state(0, torch.from_numpy(u), torch.from_numpy(v))
    fl.project()
    st = fl.stats()
    h = 2 * np.pi / N
    mg = F.build_mg(N)
    b = F.divergence_rhs(u, v, h)
    res = F.project_solve(mg, b, np.zeros((N, N)), tol=1e-6)
    assert int(st["iters"][0]) == res.iters and int(st["status"][0]) == 1
    phi = fl.phi[0].cpu().numpy()
    assert rel(phi - phi.mean(), res.phi - res.phi.mean()) < 1e-6
    ug = fl.u[0].cpu().numpy()
    ur = (u.astype(np.float64) - (res.phi - np.roll(res.phi, 1, 1)) * np.float64(np.float32(1 / h)))
    assert rel(ug, ur) < 1e-6
    assert float(st["rel"][0]) <= 1e-6


def test_laminar_state_stays_put(dev):
    import torch
    N, nu = 128, 0.01
    u0, v0 = F.laminar_state(N, nu)
    fl = _flow(dev, N)
    fl.set_state(0, torch.from_numpy(u0), torch.from_numpy(v0))
    fl.step(20)
    u, v, _ = (t.cpu().numpy() for t in fl.state(0))
    assert np.abs(u - u0).max() < 5e-5 * np.abs(u0).max()
    assert np.abs(v).max() < 1e-5 * np.abs(u0).max()


def test_flow_is_the_same_alone_and_in_an_ensemble(dev):
    import torch
    N = 64
    fl = _flow(dev, N, nflows=3)
    for q in range(3):
        u, v = F.random_modes_ic(N, 10 + q)
        fl.set_state(q, torch.from_numpy(u), torch.from_numpy(v))
    fl.step(10)
    one = _flow(dev, N, nflows=1, team=1)
    u, v = F.random_modes_ic(N, 11)
    one.set_state(0, torch.from_numpy(u), torch.from_numpy(v))
    one.step(10)
    for a, b in zip(one.state(0), fl.state(1)):
        assert bool((a == b).all())
    assert (one.stats()[0].tolist() == fl.stats()[1].tolist())


def test_bad_grids_are_rejected(dev):
    from hfm.flow2d import KolmogorovFlow
    with pytest.raises(ValueError):
        KolmogorovFlow(N=96, device=dev)
    with pytest.raises(ValueError):
        KolmogorovFlow(N=256, nlev=9, device=dev)


@pytest.mark.parametrize("team", [4, 32])
def test_team_flow_matches_oracle(dev, team):
    # one flow spread over a team of workgroups (cooperative launch, counter
    # barriers): the oracle's steps with the same substep and CG counts, and
    # the single-workgroup flow to rounding (reductions summed per workgroup)
    import torch
    N, steps, seed = 256, 3, 1
    fl = _flow(dev, N, team=team)
    u, v = F.random_modes_ic(N, seed)
    fl.set_state(0, torch.from_numpy(u), torch.from_numpy(v))
    fl.step(steps)
    st = fl.stats()
    ref, nsubs = _oracle_run(N, seed, steps)
    assert int(st["status"][0]) == 1 and int(st["steps"][0]) == steps
    assert int(st["substeps"][0]) == sum(nsubs) and int(st["iters_sum"][0]) == sum(ref.iters)
    ug, vg, phig = (t.cpu().numpy() for t in fl.state(0))
    assert rel(ug, ref.u) < 1e-5 and rel(vg, ref.v) < 1e-5
    solo = _flow(dev, N, team=1)
    solo.set_state(0, torch.from_numpy(u), torch.from_numpy(v))
    solo.step(steps)
    for a, b in zip(fl.state(0), solo.state(0)):
        assert rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-9


@pytest.mark.parametrize("team", [1, 4])
def test_nan_velocity_stops_with_breakdown(dev, team):
    """A NaN cell stops its flow with "breakdown" before any step (the velocity
    maximum keeps the NaN), and a NaN reaching the projection is a breakdown too,
    never a converged solve; the other flow of an ensemble steps as it would alone."""
    import torch
    from hfm import _lib as L
    N = 64
    nflows = 2 if team == 1 else 1
    fl = _flow(dev, N, nflows=nflows, team=team)
    u, v = F.random_modes_ic(N, 3)
    u_bad = u.copy()
    u_bad[7, 9] = np.nan
    fl.set_state(0, torch.from_numpy(u_bad), torch.from_numpy(v))
    if nflows == 2:
        fl.set_state(1, torch.from_numpy(u), torch.from_numpy(v))
    fl.step(3)
    st = fl.stats()
    assert L.FLOW_STATUS[int(st["status"][0])] == "breakdown"
    assert int(st["steps"][0]) == 0
    if nflows == 2:
        one = _flow(dev, N, nflows=1, team=1)
        one.set_state(0, torch.from_numpy(u), torch.from_numpy(v))
        one.step(3)
        assert int(st["status"][1]) == 1 and int(st["steps"][1]) == 3
        for a, b in zip(one.state(0), fl.state(1)):
            assert bool((a == b).all())
    # projection only: the NaN reaches the right-hand side
    fl.set_state(0, torch.from_numpy(u_bad), torch.from_numpy(v))
    fl.project()
    st = fl.stats()
    assert L.FLOW_STATUS[int(st["status"][0])] == "breakdown"
    assert int(st["iters"][0]) == 0
