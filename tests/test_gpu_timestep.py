"""The urban-flow time step and its fused kernels against the oracle (GPU)."""
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


def test_fused_advection_matches_layers(dev):
    from hfm import layers
    from hfm.timestep import urban_kernels
    k = urban_kernels(0.4, 1.0, 1 / 20)
    K = (k["dif"], *k["adv"])
    rng = np.random.default_rng(0)
    sh = (6, 10, 24)
    u, v, w, a = (rng.standard_normal(sh).astype(np.float32) for _ in range(4))
    outs = layers.advect([T(x, dev) for x in (u, v, w)], [T(x, dev) for x in (u, v, w)],
                         [T(x, dev) for x in (u, v, w)], 0.5, K)
    C = lambda x, Kk: O.conv_same(x.astype(np.float64), Kk)  # noqa: E731
    for g, phi in zip(outs, (u, v, w)):
        ref = 0.5 * (((C(phi, K[0]) - u * C(phi, K[1])) - v * C(phi, K[2])) - w * C(phi, K[3])) + phi
        assert rel(H(g), ref) < 1e-6
    (ga,) = layers.advect([T(x, dev) for x in (u, v, w)], [T(a, dev)], [T(a, dev)], 1.0, K)
    ref = (((C(a, K[0]) - u * C(a, K[1])) - v * C(a, K[2])) - w * C(a, K[3])) + a
    assert rel(H(ga), ref) < 1e-6
    b = layers.conv_sum([T(x, dev) for x in (u, v, w)], k["div"], scale=-1.0, source_bc=True)
    refb = O.boundary_condition_source(-(C(u, k["div"][0]) + C(v, k["div"][1]) + C(w, k["div"][2])))
    assert rel(H(b), refb) < 1e-6
    g0 = layers.FixedConv(k["adv"][0]).subtract_from(T(u, dev), T(a, dev))
    assert rel(H(g0), u - C(a, k["adv"][0])) < 1e-6
    r = layers.FixedConv(k["poisson"]).residual(T(a, dev), T(u, dev), zero_outlet=True)
    assert rel(H(r), O.zero_outlet_plane(C(a, k["poisson"]) - u)) < 1e-6


@pytest.mark.parametrize("sh", [(5, 11, 70), (3, 1, 130), (19, 7, 66), (20, 3, 2)])
def test_divergence_ragged_rows(dev, sh):
    """The three-field divergence walker computes two rows per thread: grids
    whose row count is odd (the last pair has one row) or one row high, chunks
    ending mid-walk, and tiles wider or narrower than a wave."""
    from hfm import layers
    from hfm.timestep import urban_kernels
    k = urban_kernels(0.4, 1.0, 1 / 20)
    rng = np.random.default_rng([sh[0], sh[1], sh[2]])
    u, v, w = (rng.standard_normal(sh).astype(np.float32) for _ in range(3))
    print("divergence", sh, flush=True)
    b = layers.conv_sum([T(x, dev) for x in (u, v, w)], k["div"], scale=-1.0, source_bc=True)
    C = lambda x, Kk: O.conv_same(x.astype(np.float64), Kk)  # noqa: E731
    refb = O.boundary_condition_source(-(C(u, k["div"][0]) + C(v, k["div"][1]) + C(w, k["div"][2])))
    assert rel(H(b), refb) < 1e-6


def _urban_state(sh, seed):
    rng = np.random.default_rng(seed)
    u = (1.0 + 0.05 * rng.standard_normal(sh)).astype(np.float32)
    v, w, p = (0.05 * rng.standard_normal(sh).astype(np.float32) for _ in range(3))
    a = np.zeros(sh, np.float32); a[:4] = 1
    sig = np.zeros(sh, np.float32); sig[1:5, 6:9, 6:9] = 1e5
    return (u, v, w, p, a), sig


@pytest.mark.parametrize("coarsest,j_itr", [(1, 1), (2, 1), (1, 2)])
def test_urban_steps_match_oracle(dev, coarsest, j_itr):
    """The reference's urban step (UM/AI_HFM_v3.ipynb cell 25): pressure cycles with
    the prolonged fine correction applied unsmoothed; the divergence guard reads it."""
    from hfm.timestep import UrbanFlow
    sh = (8, 16, 16)
    o, sig = _urban_state(sh, 1)
    flow = UrbanFlow(sh, sigma=T(sig, dev), multi_itr=3, coarsest=coarsest, j_itr=j_itr)
    g = [T(x, dev) for x in o]
    for _ in range(3):
        g = flow.step(*g)
        res = O.urban_step(*o, sig, multi_itr=3, coarsest=coarsest, j_itr=j_itr)
        o = res[:5]
    for gi, oi in zip(g, o):
        assert rel(H(gi), oi) < 1e-5
    w_fine = res[5]
    assert rel(H(flow.last_w), w_fine) < 1e-5
    assert flow.diverged(limit=0.0) == bool(np.abs(w_fine).max() > 0.0)
    assert not flow.diverged()
    lim = float(np.abs(w_fine).max())
    assert flow.diverged(limit=0.5 * lim) and not flow.diverged(limit=2.0 * lim)


@pytest.mark.parametrize("kind", ["k27", "k7", "var6", "2d"])
@pytest.mark.parametrize("outlet", [False, True])
def test_prolong_correct_jacobi_is_prolong_then_correct(dev, kind, outlet):
    """hfm_prolong_correct_jacobi_f32 = hfm_prolong_nearest_f32 + hfm_correct_jacobi_f32, bitwise."""
    import torch
    from hfm import _lib as L, layers
    rng = np.random.default_rng(21)
    sh = (1, 12, 20) if kind == "2d" else (6, 10, 24)
    fz = 1 if kind == "2d" else 2
    csh = (sh[0] // fz, sh[1] // 2, sh[2] // 2)
    x, b = (T(rng.standard_normal(sh), dev) for _ in range(2))
    wc = T(rng.standard_normal(csh), dev)
    if kind == "var6":
        mats = {k: T(rng.uniform(0.5, 2.0, sh), dev) for k in layers.VAR6_DIRS}
        op = layers.AOperatorVar6(mats)
        K, d, m, tot = None, 0.0, op._arr, op.tot.data_ptr()
    else:
        Kw = O.poisson27_kernel() if kind != "k7" else O.a7_const_kernel()
        if kind == "2d":
            Kw = O.implicit_ad2d_kernel()
        K27 = layers.kernel27(Kw)
        K, d, m, tot = K27.ctypes.data_as(L.C.c_void_p), float(K27[13]), None, None
    flags = L.CORRECT_ZERO_OUTLET if outlet else 0
    s = L.stream_handle(dev)
    wf = torch.empty(sh, device=dev)
    L.call("hfm_prolong_nearest_f32", wc.data_ptr(), wf.data_ptr(), *csh, fz, 2, 2, s)
    y1, y2 = torch.empty(sh, device=dev), torch.empty(sh, device=dev)
    L.call("hfm_correct_jacobi_f32", x.data_ptr(), wf.data_ptr(), b.data_ptr(), y1.data_ptr(),
           K, d, m, tot, *sh, flags, s)
    L.call("hfm_prolong_correct_jacobi_f32", x.data_ptr(), wc.data_ptr(), b.data_ptr(),
           y2.data_ptr(), K, d, m, tot, *sh, fz, 2, 2, flags, s)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("which", ["UPWIND", "CENTRAL"])
def test_ad2d_steps_match_oracle(dev, which):
    from hfm.timestep import AdvectionDiffusion2D
    K = getattr(O, "AD2D_" + which)
    x = np.zeros((60, 70), np.float32)
    x[20:40, 25:45] = 1.0
    drv = AdvectionDiffusion2D(K)
    g = T(x[None, ..., None], dev)
    o = x.astype(np.float64)
    for _ in range(20):
        g = drv.step(g)
        o = O.ad2d_step(o, K)
    assert tuple(g.shape) == (1, 60, 70, 1)
    assert rel(H(g)[0, ..., 0], o) < 1e-6
    # mass: the kernels conserve the sum away from the zero-ghost boundary
    assert abs(H(g).sum() - x.sum()) / x.sum() < 1e-3


@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
def test_urban_projection_pressure_solve(dev, storage):
    """The projection's pressure solve reaches 1e-8 against the FP64 sparse solution."""
    import torch
    from hfm.timestep import UrbanProjection
    from hfm.workloads import log_law_profile, urban_blocks
    sh = (8, 32, 32)
    mask = urban_blocks(sh, 3, foot=(3, 6), height=(2, 5))
    sig = np.where(mask, 1e5, 0.0).astype(np.float32)
    flow = UrbanProjection(sh, sigma=T(sig, dev), inflow=log_law_profile(sh[0]), dt=0.05,
                           storage=storage)
    z = torch.zeros(sh, device=dev)
    u = torch.ones(sh, device=dev)
    state = (u, z, z.clone(), z.clone(), z.clone())
    for _ in range(2):
        state = flow.step(*state)
    info = flow.last_info
    assert int(info["status"][0]) == 1 and float(info["rel"][0]) <= 1e-8
    # FP64 check of the last pressure system on the host
    inst = flow.solver.insts[0]
    b = inst.b32.cpu().numpy().reshape(sh).astype(np.float64)
    ones = np.ones(sh, np.float32)
    op = O.harmonic_faces(ones, ones, ones)
    op.cx[:, :, 0] = 0; op.cy[:, 0, :] = 0; op.cz[0] = 0; op.byh[:] = 0; op.bzh[:] = 0
    x64 = O.solve_fp64_direct(op, b)
    x = inst.solution().cpu().numpy()
    assert rel(x, x64) < 1e-6
    # the inflow profile sits on the x faces
    assert rel(H(state[0])[2, 5, 0], log_law_profile(sh[0])[2]) < 1e-6


def _fields(sh, seed, dev):
    rng = np.random.default_rng(seed)
    u, v, w, p, a = (rng.standard_normal(sh).astype(np.float32) for _ in range(5))
    sig = rng.uniform(0.0, 2.0, sh).astype(np.float32)
    sig[:, 1:3, 1:3] = 1e5                                  # an obstacle
    return [T(x, dev) for x in (u, v, w, p, a)], T(sig, dev)


@pytest.mark.parametrize("sh", [(8, 16, 16), (3, 3, 5), (5, 3, 8), (4, 7, 70), (2, 4, 6)])
@pytest.mark.parametrize("post", [False, True])
@pytest.mark.parametrize("inflow", [False, True])
def test_gradient3_equals_separate_passes_bitwise(dev, sh, post, inflow):
    """hfm_grad3_f32: the pressure fill, the damping before the gradient (or, with
    post, after it and the in-place velocity fill through
    hfm_bc_velocity_bluff_inplace_f32) and the three gradients in one pass, bitwise
    the separate layers (shapes with 2 and 3 cells per axis: the fill's source
    rows coincide with boundary rows or with each other)."""
    import torch
    from hfm import layers
    from hfm.timestep import urban_kernels
    print(f"grad3 {sh} post={post} inflow={inflow}", flush=True)
    Kg = urban_kernels(0.4, 1.0, 1 / 20)["adv"]
    (u, v, w, p, _), sig = _fields(sh, 5, dev)
    prof = T(np.linspace(0.5, 1.5, sh[0]), dev) if inflow else None
    dt = 0.4
    pb = layers.boundary_condition_pressure(p)
    if post:
        ur, vr, wr = (layers.FixedConv(K).subtract_from(f, pb) for K, f in zip(Kg, (u, v, w)))
        layers.boundary_condition_velocity_(ur, vr, wr, ub=1.0, inflow=prof)
        ur, vr, wr = layers.bluff_body(ur, vr, wr, None, sig, dt=dt)
    else:
        ub_, vb_, wb_ = layers.bluff_body(u, v, w, None, sig, dt=dt)
        ur, vr, wr = (layers.FixedConv(K).subtract_from(f, pb) for K, f in zip(Kg, (ub_, vb_, wb_)))
    uf, vf, wf = u.clone(), v.clone(), w.clone()
    uo, vo, wo, pf = layers.gradient3(p, (uf, vf, wf), Kg, sig, dt, p_bc=True, pre_bluff=not post,
                                      post_bluff=post, out=(uf, vf, wf))
    if post:
        layers.boundary_condition_velocity_bluff_(uo, vo, wo, sig, dt, ub=1.0, inflow=prof)
    torch.cuda.synchronize()
    assert torch.equal(pf, pb)
    for got, want in zip((uo, vo, wo), (ur, vr, wr)):
        assert torch.equal(got, want)


@pytest.mark.parametrize("indicator,nbluff", [(True, 2), (True, 0), (False, 1)])
def test_advect_post_equals_separate_passes_bitwise(dev, indicator, nbluff):
    """hfm_advect27_post_f32: the scalar's advection, its face fill and the dampings
    in one pass, bitwise the separate hfm_advect27_f32, hfm_bc_f32 and
    hfm_bluff_body_f32 calls."""
    import torch
    from hfm import layers
    from hfm.timestep import urban_kernels
    k = urban_kernels(0.4, 1.0, 1 / 20)
    K = (k["dif"], *k["adv"])
    (u, v, w, _, a), sig = _fields((6, 10, 70), 6, dev)
    (ref,) = layers.advect((u, v, w), (a,), (a,), 1.0, K)
    if indicator:
        ref = layers.boundary_condition_indicator(ref)
    for _ in range(nbluff):
        (ref,) = layers.bluff_body(ref, None, None, None, sig, dt=0.4)
    (got,) = layers.advect((u, v, w), (a,), (a,), 1.0, K, indicator=indicator, sigma=sig, dt=0.4,
                           nbluff=nbluff)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)


def test_bluff_fields_grouped_bitwise(dev):
    """hfm_bluff_body_f32 damps up to three fields per launch (sigma read once):
    bitwise one field per call, out of place and in place; with an output that
    is another field's input, the fields keep their one-by-one order."""
    import torch
    from hfm import _lib as L
    (u, v, w, a, _), sig = _fields((6, 10, 70), 9, dev)
    fields = [u, v, w, a]
    singles = [torch.empty_like(f) for f in fields]
    for f, o in zip(fields, singles):
        L.call("hfm_bluff_body_f32", L.ptr_array([f.data_ptr()]), L.ptr_array([o.data_ptr()]), 1,
               sig.data_ptr(), 0.4, sig.numel(), L.stream_handle(dev))
    outs = [torch.empty_like(f) for f in fields]
    L.call("hfm_bluff_body_f32", L.ptr_array([f.data_ptr() for f in fields]),
           L.ptr_array([o.data_ptr() for o in outs]), 4, sig.data_ptr(), 0.4, sig.numel(),
           L.stream_handle(dev))
    inplace = [f.clone() for f in fields]
    L.call("hfm_bluff_body_f32", L.ptr_array([f.data_ptr() for f in inplace]),
           L.ptr_array([f.data_ptr() for f in inplace]), 4, sig.data_ptr(), 0.4, sig.numel(),
           L.stream_handle(dev))
    # out[0] is in[1]: field 1 is damped after field 0 wrote it (twice)
    x0, x1 = u.clone(), v.clone()
    o1 = torch.empty_like(v)
    L.call("hfm_bluff_body_f32", L.ptr_array([x0.data_ptr(), x1.data_ptr()]),
           L.ptr_array([x1.data_ptr(), o1.data_ptr()]), 2, sig.data_ptr(), 0.4, sig.numel(),
           L.stream_handle(dev))
    twice = torch.empty_like(u)
    L.call("hfm_bluff_body_f32", L.ptr_array([singles[0].data_ptr()]),
           L.ptr_array([twice.data_ptr()]), 1, sig.data_ptr(), 0.4, sig.numel(),
           L.stream_handle(dev))
    torch.cuda.synchronize()
    for got in (outs, inplace):
        for g_, r_ in zip(got, singles):
            assert torch.equal(g_, r_)
    assert torch.equal(x1, singles[0])
    assert torch.equal(o1, twice)


@pytest.mark.parametrize("dissipative", [False, True])
@pytest.mark.parametrize("bases_shared", [False, True])
def test_advect_shared_velocities_bitwise(dev, dissipative, bases_shared):
    """The momentum stages pass the advected fields as the velocities (and, in
    the predictor, as the bases): the kernel then takes the cell's velocities and
    bases from its windows instead of loading them.  Bitwise the same launch with
    distinct copies of those fields."""
    import torch
    from hfm import layers
    from hfm.timestep import urban_kernels
    k = urban_kernels(0.4, 1.0, 1 / 20)
    (u, v, w, a, _), sig = _fields((20, 10, 70), 8, dev)
    phis = (u, v, w)
    bases = phis if bases_shared else (a, a.clone(), a.clone())
    copies = tuple(t.clone() for t in phis)
    bcopies = tuple(t.clone() for t in bases)
    if dissipative:
        dis = (a, sig, a)
        got = layers.advect_dissipative(phis, phis, bases, dis, 0.5, k["adv"])
        ref = layers.advect_dissipative(copies, phis, bcopies, dis, 0.5, k["adv"])
    else:
        K = (k["dif"], *k["adv"])
        got = layers.advect(phis, phis, bases, 0.5, K)
        ref = layers.advect(copies, phis, bcopies, 0.5, K)
    torch.cuda.synchronize()
    for g_, r_ in zip(got, ref):
        assert torch.equal(g_, r_)


@pytest.mark.parametrize("driver", ["urban", "inhale"])
def test_donated_step_equals_step_bitwise(dev, driver):
    """step(donate=True) fills the input velocities in place (the caller gives
    them up): bitwise the copying step, over two steps; the donated inputs hold
    the filled fields afterwards."""
    import torch
    from hfm.timestep import InhaleFlow, UrbanFlow
    sh = (8, 16, 32)
    (u, v, w, p, a), sig = _fields(sh, 12, dev)
    cls = UrbanFlow if driver == "urban" else InhaleFlow
    f1, f2 = cls(sh, sigma=sig, device=dev), cls(sh, sigma=sig, device=dev)
    ref = (u, v, w, p, a)
    got = tuple(t.clone() for t in ref)
    for _ in range(2):
        ref = f1.step(*ref)
        got = f2.step(*got, donate=True)
    torch.cuda.synchronize()
    for g_, r_ in zip(got, ref):
        assert torch.equal(g_, r_)
    # the caller's fields are untouched without donate
    assert torch.equal(u, _fields(sh, 12, dev)[0][0])


@pytest.mark.parametrize("driver", ["urban", "urban_noscalar", "inhale", "projection"])
def test_fused_step_equals_layer_step_bitwise(dev, driver):
    """Two time steps of each driver with the fused passes against the same step
    layer by layer (the reference's order), bitwise: u, v, w, p and the scalar."""
    import torch
    from hfm.timestep import InhaleFlow, UrbanFlow, UrbanProjection
    print(f"fused step {driver}", flush=True)
    sh = (8, 16, 32)
    (u, v, w, p, a), sig = _fields(sh, 7, dev)
    u = u * 0.05 + 1.0
    v, w, p = v * 0.05, w * 0.05, p * 0.05
    a = (a > 0).float()
    sig = torch.where(sig > 1e4, sig, torch.zeros_like(sig))
    mk = {"urban": lambda f: UrbanFlow(sh, sigma=sig, multi_itr=2, fused=f),
          "urban_noscalar": lambda f: UrbanFlow(sh, sigma=sig, multi_itr=2, lscalar=False,
                                                fused=f),
          "inhale": lambda f: InhaleFlow(sh, sigma=sig, multi_itr=2, lscalar=True, fused=f),
          "projection": lambda f: UrbanProjection(sh, sigma=sig, inflow=np.linspace(0.5, 1.0, 8),
                                                  dt=0.05, storage="f32", tune=False,
                                                  fused=f)}[driver]
    outs = []
    for fused in (False, True):
        flow = mk(fused)
        g = [t.clone() for t in (u, v, w, p, a)]
        for _ in range(2):
            g = flow.step(*g)
        torch.cuda.synchronize()
        outs.append(g)
    for x, y in zip(*outs):
        assert torch.equal(x, y)
