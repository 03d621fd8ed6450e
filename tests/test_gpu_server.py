"""The continuous-batching request loop (GPU)."""
import numpy as np
import pytest

from oracle import hfm_oracle as O

pytestmark = pytest.mark.gpu


def _fields(dev, shape, seed, sig=1.0, nd=1):
    import torch
    sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, nd)
    return [torch.from_numpy(a).to(dev) for a in (sx, sy, sz, b)]


REQS = [((32, 32, 32), 1, 0.7, 1), ((48, 48, 48), 2, 1.8, 2), ((40, 24, 32), 3, 1.0, 3),
        ((64, 64, 64), 4, 1.2, 1), ((16, 40, 24), 5, 0.5, 4)]


def test_result_is_the_same_alone_and_in_the_mix(dev):
    from hfm.mg import Params
    from hfm.server import SolveServer
    params = Params(storage="bf16", tol=1e-8)
    # alone
    alone = {}
    for rid in (2, 4):
        srv = SolveServer(params, hbm_budget=8e9, capacity=16, device=dev)
        srv.submit(rid, *_fields(dev, *REQS[rid][:2], *REQS[rid][2:]))
        (r,) = srv.drain()
        assert r.status == "converged"
        alone[rid] = (r.iters, r.x.clone())
    # mixed: two requests start, the others arrive while they iterate
    srv = SolveServer(params, hbm_budget=8e9, capacity=16, iters_per_epoch=3, device=dev)
    for rid in (0, 1):
        srv.submit(rid, *_fields(dev, *REQS[rid][:2], *REQS[rid][2:]))
    results = srv.step()
    for rid in (2, 3):
        srv.submit(rid, *_fields(dev, *REQS[rid][:2], *REQS[rid][2:]))
    results += srv.step()
    srv.submit(4, *_fields(dev, *REQS[4][:2], *REQS[4][2:]))
    results += srv.drain()
    by = {r.rid: r for r in results}
    assert sorted(by) == [0, 1, 2, 3, 4]
    for rid in (2, 4):
        assert by[rid].iters == alone[rid][0]
        assert bool((by[rid].x == alone[rid][1]).all())
    for rid, (shape, seed, sig, nd) in enumerate(REQS):
        assert by[rid].status == "converged" and by[rid].rel_residual <= 1e-8
        # every result solves its own problem (FP64 residual on the host)
        sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, nd)
        op = O.harmonic_faces(sx, sy, sz)
        x = by[rid].x.cpu().numpy()
        r = b.astype(np.float64) - O.face_apply(op.astype("f32"), x)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8


def test_admission_respects_the_hbm_budget(dev):
    from hfm.mg import Instance, Params
    from hfm.server import SolveServer
    params = Params(storage="bf16", tol=1e-8)
    one = Instance.estimate_bytes((48, 48, 48), params)
    srv = SolveServer(params, hbm_budget=2.5 * one, capacity=16, device=dev)
    for rid in range(5):
        srv.submit(rid, *_fields(dev, (48, 48, 48), 10 + rid))
    res = []
    while srv.pending():
        assert len(srv.resident) <= 2
        res += srv.step()
        assert srv.used <= srv.budget
    assert sorted(r.rid for r in res) == list(range(5))
    assert all(r.status == "converged" for r in res)
    with pytest.raises(ValueError):
        srv.submit(99, *_fields(dev, (96, 96, 96), 1))


def test_failed_request_fails_alone(dev):
    """A request with a NaN right-hand side, and one with a NaN conductivity, stop with
    "breakdown" on their own; the requests beside them in the mix converge bitwise as
    they do alone (UM/AI_HFM_v3.ipynb:2625-2627 is the reference's only guard: it ends
    the whole run)."""
    from hfm.mg import Params
    from hfm.server import SolveServer
    params = Params(storage="bf16", tol=1e-8)
    alone = {}
    for rid in (0, 3):
        srv = SolveServer(params, hbm_budget=8e9, capacity=16, device=dev)
        srv.submit(rid, *_fields(dev, *REQS[rid][:2], *REQS[rid][2:]))
        (r,) = srv.drain()
        alone[rid] = (r.iters, r.x.clone())
    srv = SolveServer(params, hbm_budget=8e9, capacity=16, iters_per_epoch=3, device=dev)
    srv.submit(0, *_fields(dev, *REQS[0][:2], *REQS[0][2:]))
    sx, sy, sz, b = _fields(dev, *REQS[1][:2], *REQS[1][2:])
    b[5, 6, 7] = float("nan")
    srv.submit(10, sx, sy, sz, b)
    sx, sy, sz, b = _fields(dev, *REQS[2][:2], *REQS[2][2:])
    sy[3, 4, 5] = float("nan")
    srv.submit(11, sx, sy, sz, b)
    srv.submit(3, *_fields(dev, *REQS[3][:2], *REQS[3][2:]))
    by = {r.rid: r for r in srv.drain()}
    assert sorted(by) == [0, 3, 10, 11]
    assert by[10].status == "breakdown" and by[11].status == "breakdown"
    for rid in (0, 3):
        assert by[rid].status == "converged"
        assert by[rid].iters == alone[rid][0]
        assert bool((by[rid].x == alone[rid][1]).all())


def test_solve_log_has_one_line_per_request_and_instance(dev):
    """The per-solve JSON lines of a server and of a batch: one per finished
    request (with its id and times) or instance, matching the results."""
    import io
    import json
    from hfm.mg import MGBatch, Params
    from hfm.server import SolveServer
    from hfm.solvelog import SolveLog
    buf = io.StringIO()
    srv = SolveServer(Params(storage="bf16", tol=1e-8), hbm_budget=8e9, capacity=16,
                      device=dev, log=SolveLog(buf))
    for rid in (0, 2, 4):
        srv.submit(rid, *_fields(dev, *REQS[rid][:2], *REQS[rid][2:]))
    res = {r.rid: r for r in srv.drain()}
    recs = [json.loads(x) for x in buf.getvalue().splitlines()]
    assert sorted(r["rid"] for r in recs) == [0, 2, 4]
    for r in recs:
        assert r["iters"] == res[r["rid"]].iters and r["status"] == "converged"
        assert r["shape"] == list(REQS[r["rid"]][0]) and r["alg_bytes"] > 0
        assert r["t_done"] >= r["t_admitted"] >= r["t_arrival"]
    buf = io.StringIO()
    mg = MGBatch([REQS[1][0], REQS[3][0]], storage="bf16", device=dev)
    for q, rid in enumerate((1, 3)):
        sx, sy, sz, b = O.make_anisotropic_fields(*REQS[rid][:2], *REQS[rid][2:])
        mg.set_problem(q, sx, sy, sz, b)
    mg.setup()
    info = mg.solve(log=SolveLog(buf))
    recs = [json.loads(x) for x in buf.getvalue().splitlines()]
    assert [r["iters"] for r in recs] == [int(v) for v in info["iter"]]
    assert all(r["mode"] == "graphs" and r["batch_instances"] == 2 for r in recs)
