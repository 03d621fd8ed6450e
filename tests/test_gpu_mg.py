"""Parity of the config-mode multigrid kernels and solver with the oracle (GPU)."""
import numpy as np
import pytest

from oracle import hfm_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _batch(dev, specs, storage, **kw):
    from hfm.mg import MGBatch
    shapes = [s[0] for s in specs]
    mg = MGBatch(shapes, storage=storage, device=dev, **kw)
    probs = []
    for i, (shape, seed, sig, nd) in enumerate(specs):
        sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, nd)
        mg.set_problem(i, sx, sy, sz, b)
        probs.append((O.harmonic_faces(sx, sy, sz), b))
    mg.setup()
    return mg, probs


def _faces_of(mg, i, lev):
    inst = mg.insts[i]
    D, H, W = inst.levels[lev]
    sizes = [D * H * W] * 3 + [D * H, D * W, H * W]
    out = []
    for o, n in zip(inst.face_off[lev], sizes):
        out.append(inst.face32[lev][o:o + n].double().cpu().numpy())
    return out


def test_faces_and_galerkin_bitwise(dev):
    specs = [((20, 13, 9), 1, 1.0, 1), ((16, 16, 16), 2, 2.0, 2)]
    mg, probs = _batch(dev, specs, "f32")
    for i, (op, _) in enumerate(probs):
        ex = op.astype("f32")
        for lev in range(len(mg.insts[i].levels)):
            got = _faces_of(mg, i, lev)
            for g, e in zip(got, (ex.cx, ex.cy, ex.cz, ex.bxh, ex.byh, ex.bzh)):
                np.testing.assert_array_equal(g, e.ravel())
            if lev + 1 < len(mg.insts[i].levels):
                ex = O.galerkin_coarsen(ex).astype("f32")


def test_coarse_inverse(dev):
    specs = [((8, 8, 8), 3, 1.0, 1), ((13, 7, 5), 4, 1.5, 1)]
    mg, probs = _batch(dev, specs, "f32")
    for i, (op, _) in enumerate(probs):
        h = O.build_hierarchy(op, "f32")
        n = int(np.prod(mg.insts[i].levels[-1]))
        got = mg.insts[i].ainv.cpu().numpy().reshape(n, n)
        np.testing.assert_allclose(got, h.coarse_inv, rtol=1e-10, atol=1e-12 * np.abs(h.coarse_inv).max())


@pytest.mark.parametrize("storage,outer", [("f32", None), ("bf16", None), ("bf16", "f32"),
                                           ("f16", None), ("f16", "f32")])
@pytest.mark.parametrize("shape", [(32, 32, 32), (20, 12, 36), (20, 13, 16)])
def test_solve_matches_oracle(dev, storage, outer, shape):
    # W = 32: 8-wide row kernels on the fine levels; W = 36: per-cell kernels;
    # (20, 13, 16): 8-wide rows with a partly filled last 512-cell segment of
    # the lane-interleaved FP64 vectors.
    # outer = "f32": BF16 V-cycle with z and p kept in FP32.
    spec = (shape, 7, 1.0, 1)
    mg, probs = _batch(dev, [spec], storage, tol=1e-8, maxit=100, outer=outer)
    info = mg.solve()
    op, b = probs[0]
    ref = O.fcg_solve(op, b, tol=1e-8, storage=storage, maxit=100, outer=outer)
    assert info["status"][0] == 1 and ref.status == "converged"
    assert int(info["iter"][0]) == ref.iters
    x = mg.solution(0).cpu().numpy()
    assert rel(x, ref.x) < 1e-5
    xd = O.solve_fp64_direct(op.astype("f32"), b.astype(np.float32))
    assert rel(x, xd) < 1e-6
    r = b - O.face_apply(op.astype("f32"), x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8


def test_batch_equals_alone_bitwise(dev):
    specs = [((32, 32, 32), 11, 0.5, 1), ((48, 48, 48), 12, 2.0, 3), ((24, 40, 16), 13, 1.2, 2)]
    mgb, _ = _batch(dev, specs, "bf16", tol=1e-8)
    ib = mgb.solve()
    for i, spec in enumerate(specs):
        mga, _ = _batch(dev, [spec], "bf16", tol=1e-8)
        ia = mga.solve()
        assert int(ia["iter"][0]) == int(ib["iter"][i])
        assert ia["status"][0] == ib["status"][i] == 1
        assert bool((mga.solution(0) == mgb.solution(i)).all())


def test_graph_and_eager_agree(dev):
    spec = ((32, 24, 40), 21, 1.0, 2)
    mg1, _ = _batch(dev, [spec], "bf16", use_graph=True)
    mg2, _ = _batch(dev, [spec], "bf16", use_graph=False)
    i1, i2 = mg1.solve(), mg2.solve()
    assert int(i1["iter"][0]) == int(i2["iter"][0])
    assert bool((mg1.solution(0) == mg2.solution(0)).all())
    # a second solve on the same resident hierarchy gives the same answer
    i3 = mg1.solve()
    assert int(i3["iter"][0]) == int(i1["iter"][0])


def test_launch_shapes_do_not_change_results(dev):
    """Every block size of the tunable launches and every coarse-cycle
    threshold gives bitwise the same solve."""
    from hfm.tune import tune
    specs = [((32, 32, 32), 31, 1.0, 1), ((48, 40, 64), 32, 1.5, 2), ((20, 12, 36), 33, 0.7, 1)]
    ref, _ = _batch(dev, specs, "bf16", coarse_cells=64)     # coarsest level only
    iref = ref.solve()
    mg, _ = _batch(dev, specs, "bf16")
    choice = tune(mg, reps=2)
    blocks = choice["blocks"]
    assert blocks and set(blocks.values()) <= {128, 256, 512}
    for cells in (choice["coarse_cells"], 64, 4096, 10 ** 6):
        mg.set_coarse_cells(cells)
        for table in (blocks, {k: 128 for k in blocks}, {k: 512 for k in blocks}):
            mg.set_blocks(table)
            info = mg.solve()
            assert list(info["iter"]) == list(iref["iter"]), (cells, table)
            for i in range(len(specs)):
                assert bool((mg.solution(i) == ref.solution(i)).all()), (cells, i)


@pytest.mark.parametrize("maxit", [3, 4])
def test_stop_at_maxit_keeps_every_solution_update(dev, maxit):
    # x64 moves on every other iteration; an odd stop leaves the last update
    # to the flush launched after the stopping CHECK.
    spec = ((32, 24, 40), 8, 1.5, 2)
    mg, probs = _batch(dev, [spec], "bf16", tol=1e-12, maxit=maxit, iters_per_graph=2)
    info = mg.solve()
    assert int(info["status"][0]) == 2 and int(info["iter"][0]) == maxit
    op, b = probs[0]
    ref = O.fcg_solve(op, b, tol=1e-12, storage="bf16", maxit=maxit)
    x = mg.solution(0).cpu().numpy()
    assert rel(x, ref.x) < 1e-5
    # the last iteration's update alone is far above that tolerance
    prev = O.fcg_solve(op, b, tol=1e-12, storage="bf16", maxit=maxit - 1)
    assert rel(ref.x, prev.x) > 1e-3


def test_stop_between_direction_and_update_applies_the_pending_update(dev):
    # iteration 0 leaves x's update pending; iteration 1 stops at ALPHA (its
    # p.Ap partials zeroed from the host: a breakdown), so the pending update
    # sits in the previous-direction buffer and the flush must take it from there
    from hfm import _lib as L
    spec = ((32, 32, 32), 5, 1.0, 1)
    mg, _ = _batch(dev, [spec], "bf16")
    plan, pool = mg.plan, mg.pool
    s = L.stream_handle(mg.device)
    mg.restart()
    plan.iteration(s, 0)
    plan.precondition(s)
    pool.finalize("beta", s)
    plan.launch_dir(1, s)
    pool.dots.zero_()
    pool.finalize("alpha", s)
    plan.launch_update(1, s)
    pool.finalize("check", s)
    plan.launch_flush(1, s)
    rec = mg.scalars()[0]
    assert int(rec["status"]) == 3 and int(rec["iter"]) == 1 and int(rec["active"]) == 0
    x_break = mg.solution(0).clone()
    one, _ = _batch(dev, [spec], "bf16", maxit=1, iters_per_graph=2)
    info = one.solve()
    assert int(info["status"][0]) == 2 and int(info["iter"][0]) == 1
    assert float(x_break.abs().max()) > 0
    assert bool((x_break == one.solution(0)).all())


def _buffer(inst, ptr):
    for t in [*inst.X, *inst.Y, *inst.B, inst.Z]:
        if t.data_ptr() == int(ptr):
            return t
    raise KeyError(ptr)


@pytest.mark.parametrize("storage,outer", [("bf16", None), ("f32", None), ("bf16", "f32"),
                                           ("f16", None)])
def test_fused_pass_pairs_equal_separate_launches(dev, storage, outer):
    """One V-cycle with the fused pass pairs (smooth + restrict, correct +
    smooth) equals the separate launches bitwise: tiles spanning the plane,
    tiles of row blocks (W = 128, 96), odd extents, several plane ranges, and
    an instance that cannot be tiled (W = 36) in the same batch.  The r.z
    partials are summed per tile instead of per chunk: equal to rounding."""
    import torch
    from hfm import _lib as L
    specs = [((33, 130, 128), 41, 1.0, 2), ((20, 13, 16), 42, 1.5, 1),
             ((70, 45, 96), 43, 0.7, 3), ((9, 7, 36), 44, 1.0, 1), ((64, 64, 48), 45, 2.0, 2)]
    print(f"fused pass pairs vs separate: storage={storage} outer={outer}", flush=True)
    a, _ = _batch(dev, specs, storage, outer=outer, fuse_levels=8,   # every level that tiles
                  fuse_min_tiles=0)
    b, _ = _batch(dev, specs, storage, outer=outer, fuse=False)
    fused_levels = {op[5] for op in a.plan.ops_down + [o for u in a.plan.ops_up for o in u]
                    if op[0] in ("smooth_restrict", "smooth2")}
    assert {0, 1, 2} <= fused_levels
    s = L.stream_handle(dev)
    for mg in (a, b):
        mg.restart()
        mg.pool.dots.fill_(float("nan"))   # every slot the cycle's dot covers is rewritten
        mg.plan.precondition(s)
    torch.cuda.synchronize()
    for ia, ib in zip(a.insts, b.insts):
        za = _buffer(ia, ia._tpl[("dir", 0)]["b"])
        zb = _buffer(ib, ib._tpl[("dir", 0)]["b"])
        assert torch.equal(za, zb), ia.shape
        for lev in range(1, len(ia.levels)):
            assert torch.equal(ia.B[lev], ib.B[lev]), (ia.shape, lev)
        da = a.pool.dots[ia.dots_off:ia.dots_off + ia.nblk]
        db = b.pool.dots[ib.dots_off:ib.dots_off + ib.nblk]
        assert bool(torch.isfinite(da).all())
        sa, sb = float(da.sum()), float(db.sum())
        assert abs(sa - sb) <= 1e-12 * abs(sb), (ia.shape, sa, sb)
    ia, ib = a.solve(), b.solve()
    assert list(ia["iter"]) == list(ib["iter"]) and all(int(v) == 1 for v in ia["status"])
    for i in range(len(specs)):
        assert rel(a.solution(i).cpu().numpy(), b.solution(i).cpu().numpy()) < 1e-6


def test_two_stream_groups_equal_one_plan_bitwise(dev):
    """The batch solved as two instance groups on two streams (each group's
    finalize touching only its own scalar slots) gives bitwise the one-plan
    solve."""
    specs = [((32, 32, 32), 51, 1.0, 1), ((48, 40, 64), 52, 1.5, 2), ((20, 12, 36), 53, 0.7, 1),
             ((40, 48, 32), 54, 2.0, 3), ((24, 24, 24), 55, 1.0, 1)]
    one, _ = _batch(dev, specs, "bf16", streams=1)
    two, _ = _batch(dev, specs, "bf16", streams=2)
    assert two._groups is not None
    i1, i2 = one.solve(), two.solve()
    assert list(i1["iter"]) == list(i2["iter"]) and all(int(v) == 1 for v in i2["status"])
    for i in range(len(specs)):
        assert bool((one.solution(i) == two.solution(i)).all()), i
    i3 = two.solve()        # again, on the cached group plans and graphs
    assert list(i3["iter"]) == list(i1["iter"])


def test_batched_setup_kernels_bitwise(dev):
    """The batched faces kernel (8-cell rows when W % 8 == 0, per cell
    otherwise) gives bitwise the per-instance kernel's faces, and the batched
    storage conversion gives torch's round-to-nearest BF16 of every level."""
    import torch
    specs = [((20, 13, 16), 61, 1.0, 1), ((17, 9, 12), 62, 2.0, 2), ((32, 24, 40), 63, 0.5, 1),
             ((9, 8, 8), 64, 1.5, 1)]
    print("batched setup kernels vs per-instance faces", flush=True)
    mg, _ = _batch(dev, specs, "bf16")         # faces from hfm_mg_faces, one instance at a time
    single = [inst.face32[0].clone() for inst in mg.insts]
    for inst in mg.insts:
        inst.face32[0].zero_()                 # faces are positive: a missed write shows
    mg.build_faces()                           # the same staged conductivities, batched
    mg.setup()
    torch.cuda.synchronize()
    for inst, ref in zip(mg.insts, single):
        assert torch.equal(inst.face32[0], ref), inst.shape
        for lev in range(len(inst.levels)):
            assert torch.equal(inst.faceT[lev], inst.face32[lev].to(torch.bfloat16)), (inst.shape, lev)


@pytest.mark.parametrize("storage", ["bf16", "f32", "f16"])
def test_large_instances_against_fp64_references(dev, storage):
    """96^3 (sigma_ln 2, 3 dipoles) and 128^3 (sigma_ln 1.5) solved together in one
    batch to 1e-8, against FP64 reference solutions of the same draws
    (tests/golden/make_fp64_references.py: the oracle's FP64 V-cycle-preconditioned
    CG to 1e-13) on 8192 fixed cells each; the true FP64 residual of the whole field
    is checked on the host."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "fp64_reference_solves.npz"))
    specs = []
    for q in range(2):
        D, H, W, seed, sig, nd = gold[f"case{q}_spec"]
        specs.append(((int(D), int(H), int(W)), int(seed), float(sig), int(nd)))
    mg, probs = _batch(dev, specs, storage, tol=1e-8, maxit=200)
    info = mg.solve()
    for q, (op, b) in enumerate(probs):
        assert int(info["status"][q]) == 1
        x = mg.solution(q).cpu().numpy()
        xs = x.ravel()[gold[f"case{q}_idx"]]
        ref = gold[f"case{q}_x"]
        err = np.linalg.norm(xs - ref) / np.linalg.norm(ref)
        assert err < 1e-6, err
        assert abs(np.linalg.norm(x) / gold[f"case{q}_norm"][0] - 1.0) < 1e-6
        r = b - O.face_apply(op.astype("f32"), x)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8


@pytest.mark.parametrize("shape,storage,tol,alpha", [
    ((32, 32, 32), "f32", 1e-6, 1.0),    # the plain V(2,2) method of the 64^3 configuration
    ((64, 64, 64), "f32", 1e-6, 1.6),
    ((48, 40, 32), "bf16", 1e-8, 1.6),   # FP64 defect correction around BF16 V-cycles
])
def test_defect_correction_matches_oracle(dev, shape, storage, tol, alpha):
    """method="refine": x += V(r), r -= A V(r) in FP64 around storage-precision
    V(2,2) cycles, against the oracle's solve_refined (one V-cycle per outer
    step): the same number of cycles, the same solution."""
    spec = (shape, 0, 1.0, 1)
    mg, probs = _batch(dev, [spec], storage, tol=tol, maxit=400, alpha=alpha, method="refine")
    info = mg.solve()
    op, b = probs[0]
    ref = O.solve_refined(op, b, tol=tol, storage=storage, inner_cycles=1, alpha=alpha,
                          max_outer=400)
    assert int(info["status"][0]) == 1 and ref.converged
    assert int(info["iter"][0]) == ref.outer_iters
    x = mg.solution(0).cpu().numpy()
    assert rel(x, ref.x) < 1e-5
    r = b - O.face_apply(op.astype("f32"), x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= tol


def test_every_tile_plane_choice_gives_the_same_cycle(dev):
    """The fused pass pairs with 8, 16, 32 or 64 planes per tile: z and every
    coarse right-hand side bitwise the same, the r.z partials' sum to rounding
    (they are summed per tile), the solves the same iterations."""
    import torch
    from hfm import _lib as L
    specs = [((64, 64, 64), 61, 1.0, 2), ((48, 96, 96), 62, 1.5, 1), ((33, 30, 32), 63, 0.7, 3)]
    s = L.stream_handle(dev)
    runs = {}
    for z in (8, 16, 32, 64):
        mg, _ = _batch(dev, specs, "bf16", fuse_min_tiles=0, tile_planes=z)
        assert any(op[0] == "smooth2" for u in mg.plan.ops_up for op in u)
        mg.restart()
        mg.plan.precondition(s)
        torch.cuda.synchronize()
        zs = [_buffer(i, i._tpl[("dir", 0)]["b"]).clone() for i in mg.insts]
        bs = [[i.B[lev].clone() for lev in range(1, len(i.levels))] for i in mg.insts]
        dsum = [float(mg.pool.dots[i.dots_off:i.dots_off + i.nblk].sum()) for i in mg.insts]
        info = mg.solve()
        runs[z] = (zs, bs, dsum, list(info["iter"]), [mg.solution(q).clone() for q in range(3)])
    ref = runs[32]
    for z, (zs, bs, dsum, its, xs) in runs.items():
        for a, b in zip(zs, ref[0]):
            assert torch.equal(a, b), z
        for la, lb in zip(bs, ref[1]):
            for a, b in zip(la, lb):
                assert torch.equal(a, b), z
        for a, b in zip(dsum, ref[2]):
            assert abs(a - b) <= 1e-12 * abs(b)
        assert its == ref[3]
        for a, b in zip(xs, ref[4]):
            assert rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-9


def test_every_schedule_converges_to_the_true_residual(dev):
    """tune_schedule's candidates (nu1, nu2, alpha) each reach 1e-8 on the FP64
    true residual of every instance (the tuner only ranks them by time)."""
    from hfm.tune import SCHEDULES, tune_schedule
    specs = [((32, 32, 32), 71, 2.0, 2), ((48, 40, 64), 72, 1.0, 1), ((20, 12, 36), 73, 0.5, 3)]
    mg, probs = _batch(dev, specs, "bf16", tol=1e-8, maxit=300)
    for sch in SCHEDULES:
        mg.set_schedule(*sch)
        info = mg.solve()
        assert bool((info["status"] == 1).all()), sch
        for q, (op, b) in enumerate(probs):
            x = mg.solution(q).cpu().numpy()
            r = b - O.face_apply(op.astype("f32"), x)
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-8, (sch, q)
    best = tune_schedule(mg)
    assert tuple(best) in {tuple(s) for s in SCHEDULES}
    assert (mg.params.nu1, mg.params.nu2, mg.params.alpha) == tuple(best)


@pytest.mark.parametrize("storage,outer,method,fuse", [
    ("bf16", None, "fcg", True), ("f32", None, "fcg", True), ("bf16", "f32", "fcg", True),
    ("bf16", None, "refine", False), ("bf16", None, "fcg", False), ("f16", None, "fcg", True)])
def test_persistent_solve_equals_the_launch_plan_bitwise(dev, storage, outer, method, fuse):
    """The persistent launch (one hfm_mg_persistent program per iters_per_graph
    iterations, per-instance dependency counters) gives bitwise the solutions,
    iteration counts and statuses of the graph-replayed launches; instances
    stop at different iterations, so the programs shrink on the way."""
    print(f"persistent vs launches: {storage} {outer} {method} fuse={fuse}", flush=True)
    specs = [((64, 64, 64), 81, 1.0, 2), ((48, 96, 32), 82, 2.0, 1), ((20, 12, 36), 83, 0.7, 3),
             ((32, 32, 32), 84, 0.5, 1), ((96, 40, 48), 85, 1.5, 4), ((9, 7, 16), 86, 1.0, 1)]
    kw = dict(tol=1e-8, maxit=300, outer=outer, method=method, fuse=fuse, fuse_min_tiles=0)
    if method == "refine":
        kw["alpha"] = 1.0 if storage == "f32" else 1.6
    mg, _ = _batch(dev, specs, storage, **kw)
    ref = mg.solve()
    xr = [mg.solution(q).clone() for q in range(len(specs))]
    got = mg.solve(persistent=True)
    assert list(got["iter"]) == list(ref["iter"]) and list(got["status"]) == list(ref["status"])
    assert len(set(int(v) for v in ref["iter"])) > 1
    for q in range(len(specs)):
        assert bool((mg.solution(q) == xr[q]).all()), q
    again = mg.solve(persistent=True)       # the cached programs
    assert list(again["iter"]) == list(ref["iter"])
    for q in range(len(specs)):
        assert bool((mg.solution(q) == xr[q]).all()), q


def test_zero_right_hand_side_stops_at_once_and_alone(dev):
    """An instance with b = 0 converges at iteration 0 with x = 0 exactly; the
    instances beside it are bitwise what they are without it."""
    import torch
    specs = [((32, 32, 32), 21, 1.0, 1), ((24, 40, 16), 22, 1.5, 2)]
    alone, _ = _batch(dev, specs, "bf16", tol=1e-8, fuse=False)
    ia = alone.solve()
    from hfm.mg import MGBatch
    mg = MGBatch([specs[0][0], (40, 24, 32), specs[1][0]], storage="bf16", device=dev, fuse=False)
    for q, spec in zip((0, 2), specs):
        sx, sy, sz, b = O.make_anisotropic_fields(*spec)
        mg.set_problem(q, sx, sy, sz, b)
    sx, sy, sz, _ = O.make_anisotropic_fields((40, 24, 32), 23, 1.0, 1)
    mg.set_problem(1, sx, sy, sz, np.zeros((40, 24, 32), np.float32))
    mg.setup()
    info = mg.solve()
    assert int(info["status"][1]) == 1 and int(info["iter"][1]) == 0
    assert float(info["rel"][1]) == 0.0
    assert bool((mg.solution(1) == 0).all())
    for q, p in ((0, 0), (2, 1)):
        assert int(info["iter"][q]) == int(ia["iter"][p])
        assert torch.equal(mg.solution(q), alone.solution(p))


@pytest.mark.parametrize("shape", [(5, 64, 64), (8, 8, 256), (128, 16, 9), (6, 6, 6)])
def test_flat_and_elongated_grids_match_oracle(dev, shape):
    """Boxes far from cubic (a 5-plane slab, a long bar, odd extents) and the
    smallest grid with two levels: the oracle's iteration count and its FP64
    direct solution."""
    spec = (shape, 31, 1.0, 1)
    mg, probs = _batch(dev, [spec], "bf16", tol=1e-8, maxit=200)
    info = mg.solve()
    op, b = probs[0]
    ref = O.fcg_solve(op, b, tol=1e-8, storage="bf16", maxit=200)
    assert info["status"][0] == 1 and ref.status == "converged"
    assert int(info["iter"][0]) == ref.iters
    x = mg.solution(0).cpu().numpy()
    xd = O.solve_fp64_direct(op.astype("f32"), b.astype(np.float32))
    assert rel(x, xd) < 1e-6


def test_largest_single_grid_reaches_the_true_residual(dev):
    """A 512^3 instance (134 M cells, the largest edge the fused tiles and the
    32-bit cell indices are sized for) to 1e-8, checked by an FP64 residual of
    the stored FP32 faces computed with torch array ops on the GPU."""
    import torch
    from hfm.mg import MGBatch
    from hfm.workloads import host_fields
    shape = (512, 512, 512)
    print("512^3 solve", flush=True)
    sx, sy, sz, b = host_fields(shape, 1.0, 2, 5)
    mg = MGBatch([shape], storage="bf16", tol=1e-8, maxit=100, device=dev)
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    mg.set_problem(0, T(sx), T(sy), T(sz), T(b))
    del sx, sy, sz
    mg.setup()
    info = mg.solve()
    assert int(info["status"][0]) == 1
    inst = mg.insts[0]
    D, H, W = shape
    N = D * H * W
    f, o = inst.face32[0], inst.face_off[0]
    cx, cy, cz = (f[o[q]:o[q] + N].view(D, H, W).double() for q in range(3))
    bxh = f[o[3]:o[3] + D * H].view(D, H).double()
    byh = f[o[4]:o[4] + D * W].view(D, W).double()
    bzh = f[o[5]:o[5] + H * W].view(H, W).double()
    x = inst.solution().double()
    ax = (cx + torch.cat([cx[:, :, 1:], bxh[:, :, None]], 2) + cy +
          torch.cat([cy[:, 1:, :], byh[:, None, :]], 1) + cz + torch.cat([cz[1:], bzh[None]], 0)) * x
    ax[:, :, 1:] -= cx[:, :, 1:] * x[:, :, :-1]
    ax[:, :, :-1] -= cx[:, :, 1:] * x[:, :, 1:]
    ax[:, 1:, :] -= cy[:, 1:, :] * x[:, :-1, :]
    ax[:, :-1, :] -= cy[:, 1:, :] * x[:, 1:, :]
    ax[1:] -= cz[1:] * x[:-1]
    ax[:-1] -= cz[1:] * x[1:]
    bt = inst.b32.view(D, H, W).double()
    relres = float((bt - ax).norm() / bt.norm())
    assert relres <= 1e-8, relres
