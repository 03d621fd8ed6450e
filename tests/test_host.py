"""Host-side checks that need no GPU: the C ABI library and the work tables."""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "hfm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hfm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from hfm import _lib as L
    lib = L.load()
    syms = _declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)


def test_abi_struct_sizes_match_host_dtypes():
    from hfm import _lib as L
    lib = L.load()
    a, b = L.C.c_int64(), L.C.c_int64()
    assert lib.hfm_struct_sizes(L.C.byref(a), L.C.byref(b)) == 0
    assert a.value == L.TASK_DTYPE.itemsize == 240
    assert b.value == L.SCALARS_DTYPE.itemsize == 120


def test_flow2d_struct_sizes_match_host_dtypes():
    from hfm import _lib as L
    lib = L.load()
    a, b = L.C.c_int64(), L.C.c_int64()
    assert lib.hfm_flow2d_struct_sizes(L.C.byref(a), L.C.byref(b)) == 0
    assert a.value == L.FLOW_TASK_DTYPE.itemsize == 112
    assert b.value == L.FLOW_PARAMS_DTYPE.itemsize == 64


def test_flow2d_rejects_bad_parameters():
    import numpy as np
    from hfm import _lib as L
    prm = np.zeros((), dtype=L.FLOW_PARAMS_DTYPE)
    prm["h"], prm["N"], prm["nu1"], prm["nu2"], prm["maxit"], prm["max_sub"] = 0.1, 96, 2, 2, 10, 4
    prm["nlev"] = 4
    with pytest.raises(L.HfmError, match="not supported"):       # 96 is not a power of two
        L.call("hfm_flow2d_steps", 1, 1, prm.ctypes.data, 1, 1, 0, None)
    prm["N"], prm["nlev"] = 256, 2                                # coarsest 128^2 > 1024 cells
    with pytest.raises(L.HfmError, match="not supported"):
        L.call("hfm_flow2d_steps", 1, 1, prm.ctypes.data, 1, 1, 0, None)
    prm["nlev"], prm["maxit"] = 6, 0
    with pytest.raises(L.HfmError, match="bad parameters"):
        L.call("hfm_flow2d_steps", 1, 1, prm.ctypes.data, 1, 1, 0, None)
    prm["maxit"] = 10
    for team in (0, L.FLOW_TEAM_MAX + 1):                        # workgroups per flow
        prm["team"] = team
        with pytest.raises(L.HfmError, match="team"):
            L.call("hfm_flow2d_steps", 1, 1, prm.ctypes.data, 1, 1, 0, None)


def test_error_path_reports_message():
    from hfm import _lib as L
    with pytest.raises(L.HfmError, match="hfm_restrict_mean_f32"):
        # odd extent: rejected on the host before any launch
        L.call("hfm_restrict_mean_f32", 1, 2, 3, 4, 5, 2, 2, 2, None)
    with pytest.raises(L.HfmError, match="unknown storage"):
        L.call("hfm_mg_convert", 1, 2, 4, 7, None)


def test_fast_division_magic_is_exact():
    from hfm import _lib as L
    rng = np.random.default_rng(0)
    for d in [1, 2, 3, 5, 7, 48, 96, 128, 208, 255, 1000, 65535]:
        m = L.magic(d)
        ns = np.concatenate([rng.integers(0, 2 ** 31, 2000), np.arange(0, 4 * d), [2 ** 31 - 1]])
        for n in ns:
            q = int(n) if m == 0 else (int(n) * m) >> 64
            assert q == int(n) // d, (n, d)


def bt_of(lay):
    return lay.blk_task.numpy().view(np.uint16)


def _owned_by(inst, ptr, nbytes):
    bufs = [*inst.X, *inst.Y, *inst.B, *inst.P, *inst.face32, *inst.faceT, inst.x64, inst.r64,
            inst.b32, inst.sigma, inst.ainv, inst.Z]
    return any(b.data_ptr() <= ptr and ptr + nbytes <= b.data_ptr() + b.numel() * b.element_size()
               for b in bufs)


@pytest.mark.parametrize("tuned", [False, True])
def test_work_tables_cover_every_cell_once(tuned):
    torch = pytest.importorskip("torch")
    from hfm import _lib as L
    from hfm.mg import ALIGN, CHUNK, MGBatch
    shapes = [(32, 32, 32), (48, 48, 48), (13, 20, 7), (96, 40, 33)]
    mg = MGBatch(shapes, storage="bf16", device="cpu", fuse_min_tiles=0)
    if tuned:   # launch shapes as the tuner may set them
        blocks = {("sweep", "plain", 0): 512, ("sweep", "correct", 0): 128,
                  ("restrict", 0): 128, ("restrict", 1): 512, ("sweep", "zero", 1): 512}
        mg.set_blocks(blocks)
        with pytest.raises(ValueError):
            mg.set_blocks({("restrict", 0): 384})
    plan = mg.plan
    by_slot = {inst.slot: inst for inst in mg.insts}
    fl = plan.flay
    bt = fl.blk_task.numpy().view(np.uint16)
    assert bt.size == fl.nblocks
    tasks = plan.init_tasks.numpy().view(L.TASK_DTYPE)
    dot_ranges = []
    for idx, t in enumerate(tasks):
        inst = by_slot[int(t["inst"])]
        assert t["blk0"] % 8 == 0 and t["nblk"] % 8 == 0
        assert np.all(bt[t["blk0"]:t["blk0"] + t["nblk"]] == idx)
        N = int(t["D"]) * int(t["H"]) * int(t["W"])
        assert t["nblk"] * CHUNK >= N > (t["nblk"] - 8) * CHUNK
        per = t["nblk"] // 8
        assert sorted((l % 8) * per + l // 8 for l in range(t["nblk"])) == list(range(t["nblk"]))
        for name, es in (("x64", 8), ("r64", 8), ("b32", 4)):
            assert _owned_by(inst, int(t[name]), N * es)
        assert t["slot0"] == inst.dots_off and t["nblk"] == inst.nblk
        dot_ranges.append((inst.dots_off, inst.dots_off + inst.nblk))
    # dot-partial ranges of different instances never overlap and fit the pool
    dot_ranges.sort()
    for (a0, a1), (b0, b1) in zip(dot_ranges, dot_ranges[1:]):
        assert a1 <= b0
    assert dot_ranges[-1][1] <= mg.pool.dots.numel()
    ops = plan.ops_down + [op for ups in plan.ops_up for op in ups]
    assert {"smooth_restrict", "smooth2"} <= {op[0] for op in ops}
    for kind, mode, tt, lay, dot, lev, block in ops:
        assert lay.blk_task.numel() == lay.nblocks
        if kind in ("smooth_restrict", "smooth2"):
            # tiles numbered contiguously over the tasks, padded to 8 with 0xFFFF
            recs = tt.numpy().view(L.TASK_DTYPE)
            total = int(recs["nblk"].sum())
            assert lay.nblocks % 8 == 0 and total <= lay.nblocks < total + 8
            assert np.all(bt_of(lay)[total:] == 0xFFFF)
            assert list(recs["blk0"]) == list(np.concatenate([[0], np.cumsum(recs["nblk"])[:-1]]))
            assert block == 512
            for t in tt.numpy().view(L.TASK_DTYPE):
                inst = by_slot[int(t["inst"])]
                shape = inst.levels[lev]
                assert shape == (t["D"], t["H"], t["W"])
                rows, planes, ntile = L.tile(shape)
                assert (t["tile_rows"], t["tile_planes"]) == (rows, planes)
                assert t["nblk"] == ntile
                assert np.all(bt_of(lay)[t["blk0"]:t["blk0"] + t["nblk"]] == bt_of(lay)[t["blk0"]])
                N = int(np.prod(shape))
                Nc = int(t["Dc"]) * int(t["Hc"]) * int(t["Wc"])
                assert (t["Dc"], t["Hc"], t["Wc"]) == inst.levels[lev + 1]
                for f in ("xin", "b", "xout"):
                    assert _owned_by(inst, int(t[f]), N * 2)
                # the first pass's input is never overwritten by the pair
                assert t["xin"] != t["xout"] and t["b"] != t["xout"]
                coarse = "rlow" if kind == "smooth_restrict" else "ec"
                assert _owned_by(inst, int(t[coarse]), Nc * 2)
                if dot:
                    assert lev == 0 and t["nblk"] <= inst.nblk   # one dot slot per block
            continue
        key = ("restrict", lev) if kind == "restrict" else \
            ("sweep", {0: "zero", 1: "plain", 2: "correct"}[mode], lev)
        assert block == (256 if dot or not tuned else blocks.get(key, 256))
        for t in tt.numpy().view(L.TASK_DTYPE):
            inst = by_slot[int(t["inst"])]
            assert inst.levels[lev] == (t["D"], t["H"], t["W"])
            N = int(t["D"]) * int(t["H"]) * int(t["W"])
            chunk = 4 * block if kind == "restrict" else 8 * block
            cells = int(t["Dc"]) * int(t["Hc"]) * int(t["Wc"]) if kind == "restrict" else N
            assert t["nblk"] * chunk >= cells and t["nblk"] % 8 == 0 and t["blk0"] % 8 == 0
            for ptr in list(t["c"]) + [t["xin"], t["xout"], t["b"], t["ec"]]:
                if ptr:
                    assert ptr % (ALIGN * 2) == 0   # 8-element aligned for 16-byte bf16 rows
            es = 2
            for f in ("xin", "b"):
                if t[f]:
                    assert _owned_by(inst, int(t[f]), N * es)
            n_out = cells if kind == "restrict" else N
            assert _owned_by(inst, int(t["xout"]), n_out * es)
            if kind == "restrict":
                assert (t["Dc"], t["Hc"], t["Wc"]) == inst.levels[lev + 1]
            if t["ec"]:
                assert _owned_by(inst, int(t["ec"]), int(t["Dc"]) * int(t["Hc"]) * int(t["Wc"]) * es)
    for t in plan.coarse_tasks.numpy().view(L.TASK_DTYPE):
        assert t["D"] * t["H"] * t["W"] <= 64


@pytest.mark.parametrize("coarse_cells", [64, 4096, 10 ** 6])
def test_coarse_cycle_chains_cover_the_tail_of_every_hierarchy(coarse_cells):
    pytest.importorskip("torch")
    from hfm import _lib as L
    from hfm.mg import MGBatch, chain_start
    shapes = [(32, 32, 32), (48, 48, 48), (13, 20, 7), (96, 40, 33), (128, 128, 128)]
    mg = MGBatch(shapes, storage="bf16", device="cpu", coarse_cells=coarse_cells)
    plan = mg.plan
    chain = plan.chain.numpy().reshape(-1, 2)
    recs = plan.chain_recs.numpy().view(L.TASK_DTYPE)
    assert len(chain) == len(mg.insts)
    ops = plan.ops_down + [op for ups in plan.ops_up for op in ups]
    for g, inst in enumerate(mg.insts):
        lc = chain_start(inst.levels, coarse_cells)
        assert 1 <= lc <= len(inst.levels) - 1
        assert all(inst.levels[l][0] * inst.levels[l][1] * inst.levels[l][2] > coarse_cells
                   for l in range(1, lc))
        first, count = chain[g]
        assert count == len(inst.levels) - lc
        for q in range(count):
            t, l = recs[first + q], lc + q
            assert (t["D"], t["H"], t["W"]) == inst.levels[l] and t["inst"] == inst.slot
            assert t["xin"] == inst.X[l].data_ptr() and t["xout"] == inst.Y[l].data_ptr()
            assert t["b"] == inst.B[l].data_ptr()
            if l + 1 < len(inst.levels):
                assert (t["Dc"], t["Hc"], t["Wc"]) == inst.levels[l + 1]
                assert t["ec"] == inst.B[l + 1].data_ptr()
            else:
                assert t["ec"] == inst.ainv.data_ptr()
        # the per-level launches stop at the chain: none touches a level >= lc,
        # every level below it has launches
        touched = set()
        for kind, mode, tt, lay, dot, lev, block in ops:
            if inst.slot in set(int(x) for x in tt.numpy().view(L.TASK_DTYPE)["inst"]):
                touched.add(lev)
        assert touched == set(range(lc)), (touched, lc)


def test_plan_rejects_bad_shapes():
    pytest.importorskip("torch")
    from hfm.mg import MGBatch
    with pytest.raises(ValueError):
        MGBatch([(4, 4, 4)], device="cpu")
    with pytest.raises(ValueError):
        MGBatch([(32, 32, 32)], storage="f8", device="cpu")
    with pytest.raises(ValueError):   # FP16 storage pairs with FP16 or FP32 z / p only
        MGBatch([(32, 32, 32)], storage="f16", outer="bf16", device="cpu")


def _ilv_cell(n):
    """Position of cell n in a lane-interleaved FP64 vector, as the kernels'
    ilv_cell computes it (512-cell segments, lane l owns cells 8l..8l+7)."""
    seg, o = n & ~511, n & 511
    return seg + 128 * ((o & 7) >> 1) + 2 * (o >> 3) + (o & 1)


@pytest.mark.parametrize("n,width", [(512, 8), (4160, 16), (32768, 32), (2080, 7), (1, 8)])
def test_fp64_vector_interleave(n, width):
    torch = pytest.importorskip("torch")
    from hfm.mg import ILV_SEG, deinterleave, ilv_len, interleave
    x = torch.arange(n, dtype=torch.float64) + 1.0
    st = interleave(x, width)
    assert st.numel() == ilv_len(n) and ilv_len(n) % ILV_SEG == 0 and ilv_len(n) - n < ILV_SEG
    assert torch.equal(deinterleave(st, n, width), x)
    if width % 8:
        assert torch.equal(st[:n], x)
        return
    # the position every kernel uses for cell m
    pos = torch.tensor([_ilv_cell(m) for m in range(n)])
    assert torch.equal(st[pos], x)
    # lane l's run: four 16-byte pairs at segment offsets 128 q + 2 l
    for seg in range(0, ilv_len(n), ILV_SEG):
        for lane in (0, 1, 37, 63):
            for q in range(4):
                for b in range(2):
                    m = seg + 8 * lane + 2 * q + b
                    want = float(m + 1) if m < n else 0.0
                    assert float(st[seg + 128 * q + 2 * lane + b]) == want


def _bgs_meshes(shape):
    """The non-uniform mesh of ART/BGS_multigrid_v1.ipynb cell 34, on a small box:
    dx grows by 0.02 per cell away from the x centre, dy by 0.4 from the y
    centre, dz by 0.02 per level."""
    D, H, W = shape
    mx, my, mz = (np.zeros(shape) for _ in range(3))
    for i in range(W // 2):
        mx[:, :, W // 2 + i] = mx[:, :, W // 2 - 1 - i] = 1.0 + 0.02 * i
    for i in range(H // 2):
        my[:, H // 2 + i, :] = my[:, H // 2 - 1 - i, :] = 1.0 + 0.4 * i
    for i in range(D):
        mz[i] = 1.0 + 0.02 * i
    return mx, my, mz


def test_mesh_shift_follows_the_reference_loop():
    pytest.importorskip("torch")
    from hfm.layers import mesh_shift
    rng = np.random.default_rng(0)
    m = rng.uniform(0.5, 2.0, (5, 6, 7))
    for axis, n in ((2, 7), (1, 6), (0, 5)):
        ref = m.copy()
        for i in range(n - 2):          # cell 36: mesh_p1[i + 1] = mesh[i + 2]
            idx_d = [slice(None)] * 3
            idx_s = [slice(None)] * 3
            idx_d[axis], idx_s[axis] = i + 1, i + 2
            ref[tuple(idx_d)] = m[tuple(idx_s)]
        np.testing.assert_array_equal(mesh_shift(m, axis), ref)
    np.testing.assert_array_equal(mesh_shift(m[None, ..., None], 2)[0, ..., 0], mesh_shift(m, 2))


def test_nonuniform_coefficients():
    pytest.importorskip("torch")
    from hfm.layers import nonuniform_coefficients
    sh = (4, 6, 8)
    ones = np.ones(sh)
    mats = nonuniform_coefficients(ones, ones, ones, 100.0, 50.0, 25.0)
    for k, v in (("xm1", 100), ("xp1", 100), ("ym1", 50), ("yp1", 50), ("zm1", 25), ("zp1", 25)):
        assert mats[k].dtype == np.float32 and (mats[k] == v).all()
    assert (mats["tot"] == 350).all()
    mx, my, mz = _bgs_meshes(sh)
    mats = nonuniform_coefficients(mx, my, mz)
    # x = 5: dx = 1.02 there, the shifted size is the next cell's 1.04
    assert mats["xm1"][0, 0, 5] == np.float32(100 / 1.02 ** 2)
    assert mats["xp1"][0, 0, 5] == np.float32(100 / 1.04 ** 2)
    assert mats["xp1"][0, 0, 7] == np.float32(100 / 1.06 ** 2)   # boundary keeps its own
    assert mats["zp1"][1, 0, 0] == np.float32(100 / 1.04 ** 2)


def _symbolic_cycle(mg, plan, inst):
    """Run ``plan``'s V-cycle symbolically for one instance: each launch reads
    the expressions held by its input buffers and writes a new expression;
    returns the expression the direction kernel reads as z."""
    from hfm import _lib as L
    from hfm.mg import chain_start
    val = {inst.B[0].data_ptr(): "r"}
    tpl = inst._tpl
    val[int(tpl[("init",)]["xout"])] = "Z0(r)"     # first sweep, fused into the update

    def rd(ptr):
        assert ptr in val, "read of a buffer no launch wrote"
        return val[ptr]

    def rec(tt):
        for t in tt.numpy().view(L.TASK_DTYPE):
            if int(t["inst"]) == inst.slot:
                return t
        return None

    def run(op):
        kind, mode, tt, _, _, lev, _ = op
        t = rec(tt)
        if t is None:
            return
        x, b = int(t["xin"]), int(t["b"])
        if kind in ("smooth_restrict", "smooth2"):
            side = int(t["rlow"] if kind == "smooth_restrict" else t["ec"])
            assert len({x, b, side, int(t["xout"])}) == 4, "a fused pair overwrites its input"
        if kind == "sweep" and mode == 0:
            val[int(t["xout"])] = f"Z{lev}({rd(b)})"
        elif kind == "sweep" and mode == 1:
            val[int(t["xout"])] = f"P{lev}({rd(x)},{rd(b)})"
        elif kind == "sweep":
            val[int(t["xout"])] = f"C{lev}({rd(x)},{rd(int(t['ec']))},{rd(b)})"
        elif kind == "restrict":
            val[int(t["xout"])] = f"R{lev}({rd(x)},{rd(b)})"
        elif kind == "smooth_restrict":
            y = f"P{lev}({rd(x)},{rd(b)})"
            val[int(t["xout"])] = y
            val[int(t["rlow"])] = f"R{lev}({y},{rd(b)})"
        else:   # smooth2
            y = f"C{lev}({rd(x)},{rd(int(t['ec']))},{rd(b)})"
            val[int(t["xout"])] = f"P{lev}({y},{rd(b)})"

    for op in plan.ops_down:
        run(op)
    p = mg.params
    lc = chain_start(inst.levels, p.coarse_cells)
    last = len(inst.levels) - 1
    out = inst.X[lc] if lc == last or (p.nu1 + p.nu2 - 1) % 2 == 0 else inst.Y[lc]
    val[out.data_ptr()] = f"V{lc}({rd(inst.B[lc].data_ptr())})"
    for ups in plan.ops_up:
        for op in ups:
            run(op)
    return rd(int(tpl[("dir", 0)]["b"]))


@pytest.mark.parametrize("nu", [(2, 2), (1, 2), (2, 3), (3, 1), (2, 1)])
@pytest.mark.parametrize("coarse_cells,fuse_levels,min_tiles",
                         [(64, 8, 0), (4096, 8, 0), (64, 1, 0), (64, 8, 10 ** 9)])
def test_fused_plan_computes_the_same_cycle(nu, coarse_cells, fuse_levels, min_tiles):
    """With fused pass pairs every instance's V-cycle is the same composition
    of passes as with separate launches, and no fused launch writes a buffer
    it reads."""
    pytest.importorskip("torch")
    from hfm.mg import MGBatch
    shapes = [(32, 32, 32), (33, 130, 128), (13, 20, 7), (20, 13, 16), (40, 40, 36), (64, 48, 96)]
    kw = dict(storage="bf16", device="cpu", nu1=nu[0], nu2=nu[1], coarse_cells=coarse_cells)
    a = MGBatch(shapes, fuse_levels=fuse_levels, fuse_min_tiles=min_tiles, **kw)
    b = MGBatch(shapes, fuse=False, **kw)
    fused = [(op[0], op[5]) for op in a.plan.ops_down + [o for u in a.plan.ops_up for o in u]
             if op[0] in ("smooth_restrict", "smooth2")]
    kinds = {k for k, _ in fused}
    if min_tiles == 0:
        assert ("smooth_restrict" in kinds) == (nu[0] >= 2)
        assert ("smooth2" in kinds) == (nu[1] >= 2)
    else:
        # too few tiles in the batch on every level: no level runs fused
        assert not fused
    assert any(lev == 0 for _, lev in fused) == (min_tiles == 0 and (nu[0] >= 2 or nu[1] >= 2))
    assert a.pool.fuse_fine == (min_tiles == 0)
    assert a.pool.fuse_coarse == (min_tiles == 0)
    assert not {"smooth_restrict", "smooth2"} & {
        op[0] for op in b.plan.ops_down + [o for u in b.plan.ops_up for o in u]}
    for ia, ib in zip(a.insts, b.insts):
        za, zb = _symbolic_cycle(a, a.plan, ia), _symbolic_cycle(b, b.plan, ib)
        assert za == zb, (ia.shape, za, zb)


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 96, 48), (33, 130, 128), (5, 300, 512),
                                   (1, 1, 8), (70, 45, 96), (128, 128, 128), (7, 13, 16)])
def test_fused_tile_geometry(shape):
    """hfm_mg_tile: a tile plane (rows plus halo rows) fits a 512-thread block
    of 4-cell runs, row blocks are even (coarse row pairs stay in one tile),
    the tiles cover the level, and planes per tile are even and at most 32."""
    from hfm import _lib as L
    D, H, W = shape
    rows, planes, ntiles = L.tile(shape)
    q = W // 4
    ny = -(-H // rows)
    if rows >= H:
        assert ny == 1 and H * q <= 512
    else:
        assert rows % 2 == 0 and (rows + 2) * q <= 512
    assert planes % 2 == 0 and planes <= 32 + 1 and planes >= min(D, 2)
    assert ntiles == ny * -(-D // planes)
    assert (ny - 1) * rows < H <= ny * rows
    for zmax in (2, 8, 16, 64):
        r2, p2, n2 = L.tile(shape, zmax)
        assert r2 == rows and p2 % 2 == 0 and p2 <= zmax + 1 and p2 >= min(D, 2)
        assert n2 == ny * -(-D // p2)


def test_fused_entry_points_reject_bad_tables():
    from hfm import _lib as L
    # W % 8 != 0; a row too wide for a tile plus halo rows in one block
    for shape in [(8, 8, 36), (8, 8, 7), (4, 4, 1024), (4, 4, 2048)]:
        assert L.tile(shape) is None
    # tables are checked before anything touches the (here fake) pointers
    with pytest.raises(L.HfmError, match="multiple of 8"):
        L.call("hfm_mg_smooth_restrict", 16, 16, 3, 16, 1, 0.5, None)
    with pytest.raises(L.HfmError, match="multiple of 8"):
        L.call("hfm_mg_smooth2", 16, 16, 12, 16, 1, 1, 0.5, 1.6, None, None)
    with pytest.raises(L.HfmError, match="unknown storage"):
        L.call("hfm_mg_smooth2", 16, 16, 8, 16, 1, 7, 0.5, 1.6, None, None)


def test_host_fields_are_the_oracles_draw():
    """The bench workload's inputs (hfm.workloads.host_fields) are the oracle's
    make_anisotropic_fields bit for bit, so the oracle can regenerate any instance."""
    from hfm.workloads import batched_specs, host_fields, host_fields_many
    from oracle import hfm_oracle as O
    specs = [((8, 10, 12), 0.7, 2, 5), ((16, 16, 16), 1.9, 4, 99)] + \
        [(s, sig, nd, seed) for (s, sig, nd, seed) in batched_specs(3, 2024) if s[0] <= 48]
    many = host_fields_many(specs, workers=3)
    for sp, got2 in zip(specs, many):
        ref = O.make_anisotropic_fields(sp[0], sp[3], sp[1], sp[2])
        for a, b, c in zip(host_fields(*sp), ref, got2):
            assert a.dtype == np.float32
            np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(c, b)


def test_random_modes_ic_is_the_oracles():
    from hfm.workloads import random_modes_ic
    from oracle import hfm_flow2d as F
    for N, seed in ((32, 3), (64, 2024)):
        for a, b in zip(random_modes_ic(N, seed), F.random_modes_ic(N, seed)):
            np.testing.assert_array_equal(a, b)


def _simulate_persistent(prog, G):
    """Run the persistent kernel's schedule on the host: workgroup w takes, per
    phase, the items g = x + 8 e (x = w % 8) of its contiguous e-range, two at a
    time for the 256-thread phases, in program order.  An item whose slot
    differs from the half's previous one first raises the count of the
    previous slot by the number of its items, then waits for
    cnt[prev][slot] == need[prev][slot]; the phase's last counts are raised at
    its end.  Returns the per-(phase, slot) counts, or raises on a deadlock (no
    workgroup can move)."""
    from hfm import _lib as L
    rows = prog.prog.numpy().view(L.PHASE_DTYPE)
    need, prev = prog.need.numpy(), prog.prev.numpy()
    whole_kinds = {L.PH["smooth_restrict"], L.PH["smooth2"], L.PH["coarse_cycle"]}
    ops = []   # per workgroup: ("signal", phase, slot, count) and ("wait", phase, slots)
    for w in range(G):
        x, j, per = w & 7, w >> 3, G >> 3
        seq = []
        for p, r in enumerate(rows):
            n = int(r["nitems"])
            n8 = (n - x + 7) // 8 if n > x else 0
            e0, e1 = n8 * j // per, n8 * (j + 1) // per
            whole = int(r["kind"]) in whole_kinds
            nh = 1 if whole else 2
            pend, npend, okslot = [-1, -1], [0, 0], [-1, -1]
            e = e0
            while e < e1:
                gs = [x + 8 * e] + ([] if whole else [x + 8 * (e + 1)] if e + 1 < e1 else [-1])
                sl = [int(prog.item_slots[p][g]) if g >= 0 else -1 for g in gs]
                fast = all(s < 0 or (s == pend[h] and s == okslot[h]) for h, s in enumerate(sl))
                if whole or not fast:
                    for h in range(nh):
                        if sl[h] != pend[h] and npend[h] > 0:
                            seq.append(("signal", p, pend[h], npend[h]))
                    seq.append(("wait", p, [s for h, s in enumerate(sl)
                                            if s >= 0 and s != okslot[h]]))
                    for h in range(nh):
                        if sl[h] != pend[h]:
                            pend[h], npend[h] = sl[h], 0
                        okslot[h] = sl[h]
                for h, s in enumerate(sl):
                    if s >= 0:
                        npend[h] += 1
                e += 1 if whole else 2
            for h in range(nh):
                if npend[h] > 0:
                    seq.append(("signal", p, pend[h], npend[h]))
        ops.append(seq)
    cnt = np.zeros_like(need)
    pos = [0] * G
    left = sum(len(s) for s in ops)
    while left:
        moved = False
        for w in range(G):
            while pos[w] < len(ops[w]):
                op = ops[w][pos[w]]
                if op[0] == "signal":
                    cnt[op[1], op[2]] += op[3]
                else:
                    p = op[1]
                    if not all(prev[p, s] < 0 or cnt[prev[p, s], s] >= need[prev[p, s], s]
                               for s in op[2]):
                        break
                pos[w] += 1
                left -= 1
                moved = True
        if not moved:
            raise AssertionError("persistent schedule deadlocks")
    return cnt


@pytest.mark.parametrize("fuse", [True, False])
def test_persistent_program_schedule(fuse):
    """The persistent program built from a plan: every phase's predecessor per slot
    is an earlier phase, every item belongs to a slot of the plan (or is a
    padding tile), the phases follow Plan.iteration, and the kernel's static
    schedule completes every item without a deadlock for several grid sizes."""
    from hfm import _lib as L
    from hfm.mg import MGBatch, PersistentProgram
    shapes = [(64, 64, 64), (48, 96, 32), (20, 12, 36), (32, 32, 32), (9, 7, 16)]
    mg = MGBatch(shapes, storage="bf16", device="cpu", fuse=fuse, fuse_min_tiles=0)
    prog = PersistentProgram(mg.plan, 2)
    rows = prog.prog.numpy().view(L.PHASE_DTYPE)
    kinds = [int(k) for k in rows["kind"]]
    per_it = len(kinds) // 2
    assert kinds[:per_it] == kinds[per_it:]
    tail = [L.PH[k] for k in ("finalize", "dir", "finalize", "update", "finalize", "flush")]
    assert kinds[per_it - 6:per_it] == tail
    assert (L.PH["smooth_restrict"] in kinds) == fuse
    need, prev = prog.need.numpy(), prog.prev.numpy()
    for p in range(prog.nphase):
        assert np.all(prev[p] < p)
        sl = prog.item_slots[p]
        assert need[p].sum() == np.count_nonzero(sl >= 0)
        assert set(sl[sl >= 0]) <= {i.slot for i in mg.insts} | set(range(mg.pool.capacity))
    for G in (8, 16, 64, 256):
        cnt = _simulate_persistent(prog, G)
        assert np.array_equal(cnt, need)


def test_guard_bands_report_a_write_past_either_end():
    """The guard-band allocator of the GPU bounds tests (on CPU tensors): a byte
    written just before or just after a buffer is reported, an untouched one
    is not, and the buffer itself comes back zeroed."""
    import torch
    from test_gpu_guard import GUARD, GuardedAlloc
    alloc = GuardedAlloc("cpu")
    a = alloc(100, torch.float32)
    b = alloc(33, torch.bfloat16)
    assert bool((a == 0).all()) and bool((b == 0).all())
    a.fill_(7.0)
    b.fill_(1.0)
    assert alloc.damaged() == []
    raw, nb = alloc.bufs[1]
    raw[GUARD + nb] = 0
    assert alloc.damaged() == [1]
    raw[GUARD + nb] = 0xA5
    alloc.bufs[0][0][GUARD - 1] = 1
    assert alloc.damaged() == [0]


def test_solve_log_lines_and_iteration_bytes():
    """One JSON object per line with the solve's fields; the iteration's bytes
    are the per-pass figures of the kernel table (BF16: sweep 12 B/cell,
    restriction 10.25, direction 18, update 34 / 52 -> 43 on average)."""
    import io
    import json
    from hfm import _lib as L
    from hfm.mg import MGBatch, iteration_bytes
    from hfm.solvelog import SolveLog, solve_record
    mg = MGBatch([(8, 8, 8), (16, 16, 16)], storage="bf16", device="cpu")
    inst = mg.insts[0]
    assert inst.levels == [(8, 8, 8), (4, 4, 4)]
    assert iteration_bytes(inst, mg.params) == 512 * (4 * 12 + 10.25 + 18 + 43)
    rec = np.zeros(1, L.SCALARS_DTYPE)[0]
    rec["iter"], rec["status"], rec["rel"] = 7, 1, 3e-9
    buf = io.StringIO()
    log = SolveLog(buf)
    log.write(solve_record(inst, mg.params, rec, 0.5, 2, iteration_bytes(inst, mg.params),
                           rid=11))
    log.write(solve_record(mg.insts[1], mg.params, rec, 0.5, 2, 1e6))
    lines = buf.getvalue().splitlines()
    assert len(lines) == 2
    a = json.loads(lines[0])
    assert a["shape"] == [8, 8, 8] and a["levels"] == 2 and a["status"] == "converged"
    assert a["iters"] == 7 and a["rel_residual"] == 3e-9 and a["rid"] == 11
    assert a["alg_bytes"] == 7 * 512 * (4 * 12 + 10.25 + 18 + 43)
    assert abs(a["alg_GBps_share"] - a["alg_bytes"] / 0.5 / 1e9) < 1e-12
    assert json.loads(lines[1])["storage"] == "bf16"
