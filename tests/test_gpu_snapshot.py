"""Restart from snapshot files continues the time-step drivers bitwise (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _urban_state(sh, dev):
    import torch
    u = np.ones(sh, np.float32); v = np.zeros(sh, np.float32); w = np.zeros(sh, np.float32)
    p = np.zeros(sh, np.float32); a = np.zeros(sh, np.float32); a[:4] = 1
    return [torch.from_numpy(x).to(dev) for x in (u, v, w, p, a)]


@pytest.mark.parametrize("driver", ["UrbanFlow", "UrbanProjection"])
def test_urban_restart_continues_bitwise(dev, tmp_path, driver):
    import torch
    from hfm import timestep
    from hfm.snapshot import SnapshotWriter, load_restart
    sh = (8, 16, 16)
    sig = torch.zeros(sh, device=dev)
    sig[1:5, 6:9, 6:9] = 1e5
    kw = {"multi_itr": 3} if driver == "UrbanFlow" else {"dt": 0.05, "tune": False}
    mk = lambda: getattr(timestep, driver)(sh, sigma=sig, device=dev, **kw)  # noqa: E731
    straight = mk().run(*_urban_state(sh, dev), 4)
    flow = mk()
    with SnapshotWriter(str(tmp_path), n_out=2) as wr:
        flow.run(*_urban_state(sh, dev), 2, writer=wr)
    f, ctime = load_restart(str(tmp_path), 2, flow.dt, names=("u", "v", "w", "p", "alpha"),
                            device=dev, dtype=torch.float32)
    assert ctime == pytest.approx(2 * flow.dt)
    again = mk().run(*(f[n] for n in ("u", "v", "w", "p", "alpha")), 2, nrestart=2)
    for a, b in zip(straight, again):
        assert bool((a == b).all())


def test_kolmogorov_snapshots_and_restart(dev, tmp_path):
    import torch
    from hfm.flow2d import KolmogorovFlow
    from hfm.snapshot import SnapshotWriter
    from oracle import hfm_flow2d as F
    N = 64
    u0, v0 = F.random_modes_ic(N, 3)

    def fresh():
        fl = KolmogorovFlow(N=N, nlev=4, device=dev)
        fl.set_state(0, torch.from_numpy(u0), torch.from_numpy(v0))
        return fl

    ref = fresh()
    ref.step(6)
    fl = fresh()
    with SnapshotWriter(str(tmp_path), n_out=4) as wr:
        fl.run(6, writer=wr)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["p4.npy", "u4.npy", "v4.npy"]
    assert np.load(tmp_path / "p4.npy").dtype == np.float64
    for a, b in zip(fl.state(0), ref.state(0)):       # chunked launches = one launch
        assert bool((a == b).all())
    re = KolmogorovFlow(N=N, nlev=4, device=dev)
    re.restart(str(tmp_path), 4)
    re.run(2)
    for a, b in zip(re.state(0), ref.state(0)):
        assert bool((a == b).all())
