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
    assert b.value == L.SCALARS_DTYPE.itemsize == 96


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


def test_work_tables_cover_every_cell_once():
    torch = pytest.importorskip("torch")
    from hfm import _lib as L
    from hfm.mg import ALIGN, CHUNK, RCHUNK, MGBatch
    shapes = [(32, 32, 32), (48, 48, 48), (13, 20, 7), (96, 40, 33)]
    mg = MGBatch(shapes, storage="bf16", device="cpu")
    # the fine layout: blocks start at multiples of 8, cover every cell, map to the right task
    fl = mg.flay
    bt = fl.blk_task.numpy().view(np.uint16)
    assert bt.size == fl.nblocks
    tasks = mg.init_tasks.numpy().view(L.TASK_DTYPE)
    for idx, t in enumerate(tasks):
        assert t["blk0"] % 8 == 0 and t["nblk"] % 8 == 0
        assert np.all(bt[t["blk0"]:t["blk0"] + t["nblk"]] == idx)
        N = t["D"] * t["H"] * t["W"]
        assert t["nblk"] * CHUNK >= N > (t["nblk"] - 8) * CHUNK
        # XCD chunk map is a permutation of 0..nblk-1
        per = t["nblk"] // 8
        ch = sorted((l % 8) * per + l // 8 for l in range(t["nblk"]))
        assert ch == list(range(t["nblk"]))
        # pointers land inside the buffers they belong to
        for name, buf in (("x64", mg.x64), ("r64", mg.r64), ("b32", mg.b32)):
            lo = buf.data_ptr()
            assert lo <= t[name] and t[name] + N * buf.element_size() <= lo + buf.numel() * buf.element_size()
    # every sweep / restriction task's vectors stay inside its level's buffers
    for kind, mode, tt, lay, dot in mg.ops_down + [op for ups in mg.ops_up for op in ups]:
        arr = tt.numpy().view(L.TASK_DTYPE)
        for t in arr:
            lev = next(l for l in range(mg.nlev) if mg.levels[t["inst"]][l] == (t["D"], t["H"], t["W"]))
            N = t["D"] * t["H"] * t["W"]
            chunk = RCHUNK if kind == "restrict" else CHUNK
            cells = t["Dc"] * t["Hc"] * t["Wc"] if kind == "restrict" else N
            assert t["nblk"] * chunk >= cells and t["nblk"] % 8 == 0 and t["blk0"] % 8 == 0
            for ptr in list(t["c"]) + [t["xin"], t["xout"], t["b"], t["ec"]]:
                if ptr:
                    assert ptr % (ALIGN * 2) == 0   # 8-element aligned for 16-byte bf16 rows
            for f in ("xin", "xout", "b"):
                if t[f] == 0:
                    continue
                target_lev = lev + 1 if (kind == "restrict" and f == "xout") else lev
                n = N if target_lev == lev else t["Dc"] * t["Hc"] * t["Wc"]
                bufs = [mg.X[target_lev], mg.Y[target_lev], mg.B[target_lev]]
                assert any(b.data_ptr() <= t[f] and t[f] + n * b.element_size()
                           <= b.data_ptr() + b.numel() * b.element_size() for b in bufs), (kind, f)
            if kind == "restrict":
                assert (t["Dc"], t["Hc"], t["Wc"]) == tuple((v + 1) // 2 for v in (t["D"], t["H"], t["W"]))
    # coarsest tasks fit the dense-inverse kernel
    for t in mg.coarse_tasks.numpy().view(L.TASK_DTYPE):
        assert t["D"] * t["H"] * t["W"] <= 64


def test_plan_rejects_bad_shapes():
    pytest.importorskip("torch")
    from hfm.mg import MGBatch
    with pytest.raises(ValueError):
        MGBatch([(4, 4, 4)], device="cpu")
    with pytest.raises(ValueError):
        MGBatch([(32, 32, 32)], storage="f16", device="cpu")
