"""Snapshot / restart files in the reference's format (host side, CPU)."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _fields(shape, seed=0, five_d=True):
    rng = np.random.default_rng(seed)
    out = {}
    for n in ("u", "v", "w", "p"):
        a = rng.standard_normal(shape).astype(np.float32)
        out[n] = torch.from_numpy(a[None, ..., None] if five_d else a)
    return out


def test_save_data_follows_the_reference_rule(tmp_path):
    # ARTs/naiver_stokes_3D.py:488-497: when itime % n_out == 0,
    # np.save(dir + name + str(itime), arr=values[0, :, :, :, 0])
    from hfm.snapshot import save_data
    f = _fields((3, 4, 5))
    assert save_data(f, 3, 4, str(tmp_path)) == []
    assert not list(tmp_path.iterdir())
    paths = save_data(f, 3, 6, str(tmp_path))
    assert sorted(os.path.basename(p) for p in paths) == ["p6.npy", "u6.npy", "v6.npy", "w6.npy"]
    for n, t in f.items():
        a = np.load(tmp_path / f"{n}6.npy", allow_pickle=False)
        assert a.dtype == np.float32 and a.shape == (3, 4, 5)
        np.testing.assert_array_equal(a, t.numpy()[0, :, :, :, 0])


def test_writer_and_restart_round_trip(tmp_path):
    from hfm.snapshot import SnapshotWriter, load_restart
    steps = {it: _fields((4, 6, 8), seed=it, five_d=False) for it in range(1, 7)}
    with SnapshotWriter(str(tmp_path), n_out=2) as wr:
        saved = [it for it in steps if wr.maybe_save(it, **steps[it])]
    assert saved == [2, 4, 6]
    assert sorted(os.listdir(tmp_path)) == sorted(f"{n}{i}.npy" for i in (2, 4, 6) for n in "uvwp")
    for it in (2, 4, 6):
        got, ctime = load_restart(str(tmp_path), it, 0.05)
        assert ctime == pytest.approx(it * 0.05)
        for n in "uvwp":
            assert torch.equal(got[n], steps[it][n])
    # the reference casts to float32 on load; a float64 field keeps its bits otherwise
    np.save(tmp_path / "q9.npy", np.arange(6, dtype=np.float64).reshape(2, 3) / 3)
    q = load_restart(str(tmp_path), 9, 1.0, names=("q",))[0]["q"]
    assert q.dtype == torch.float64
    q32 = load_restart(str(tmp_path), 9, 1.0, names=("q",), dtype=torch.float32)[0]["q"]
    assert q32.dtype == torch.float32


def test_restart_errors(tmp_path):
    from hfm.snapshot import load_restart, save_data
    save_data(_fields((2, 3, 4)), 1, 5, str(tmp_path))
    with pytest.raises(FileNotFoundError):
        load_restart(str(tmp_path), 7, 1.0)
    np.save(tmp_path / "w5.npy", np.zeros((2, 3, 5), np.float32))
    with pytest.raises(ValueError):
        load_restart(str(tmp_path), 5, 1.0)
    # pickled payloads are never loaded
    np.save(tmp_path / "w5.npy", np.array([{"x": 1}], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):
        load_restart(str(tmp_path), 5, 1.0)


def test_writer_rejects_a_zero_period(tmp_path):
    from hfm.snapshot import SnapshotWriter
    with pytest.raises(ValueError):
        SnapshotWriter(str(tmp_path), n_out=0)


@pytest.mark.timeout(60)
def test_writer_keeps_its_pools_after_staging_errors(tmp_path, monkeypatch):
    """A snapshot that fails while staging hands its staging pool back, so later
    snapshots never block (with depth 1, a lost pool would hang the next call)."""
    from hfm import snapshot

    class FailingAlloc:
        def __getattr__(self, name):
            return getattr(torch, name)

        def empty(self, *a, **k):
            raise RuntimeError("staging buffer allocation failed")

    with snapshot.SnapshotWriter(str(tmp_path), n_out=1, depth=1) as wr:
        monkeypatch.setattr(snapshot, "torch", FailingAlloc())
        for _ in range(3):
            with pytest.raises(RuntimeError, match="staging buffer"):
                wr.maybe_save(1, **_fields((2, 3, 4), five_d=False))
        monkeypatch.setattr(snapshot, "torch", torch)
        assert wr.maybe_save(2, **_fields((2, 3, 4), five_d=False))
    assert sorted(os.listdir(tmp_path)) == [f"{n}2.npy" for n in "puvw"]
