"""Guard bands around device buffers (GPU): every buffer a solve or a layer
kernel writes sits inside a larger allocation between two 4 KiB bands of a
fixed byte pattern; after the work the bands must be untouched, so no kernel
writes past either end of the arrays it is given (ragged shapes, per-cell and
8-cell-run paths, fused tiles, the persistent kernel, the reference-mode
walkers)."""
import numpy as np
import pytest

from oracle import hfm_oracle as O

pytestmark = pytest.mark.gpu

GUARD = 4096
PATTERN = 0xA5


class GuardedAlloc:
    """alloc(numel, dtype) -> a zeroed view inside a guarded buffer."""

    def __init__(self, dev):
        self.dev, self.bufs = dev, []

    def __call__(self, n, dtype):
        import torch
        nb = int(n) * torch.empty((), dtype=dtype).element_size()
        raw = torch.full((nb + 2 * GUARD,), PATTERN, dtype=torch.uint8, device=self.dev)
        raw[GUARD:GUARD + nb].zero_()
        self.bufs.append((raw, nb))
        return raw[GUARD:GUARD + nb].view(dtype)

    def empty_like(self, t):
        return self(t.numel(), t.dtype).view(t.shape)

    def damaged(self):
        import torch
        if torch.device(self.dev).type == "cuda":
            torch.cuda.synchronize()
        bad = []
        for q, (raw, nb) in enumerate(self.bufs):
            if not bool((raw[:GUARD] == PATTERN).all()) or \
                    not bool((raw[GUARD + nb:] == PATTERN).all()):
                bad.append(q)
        return bad


SPECS = [((32, 32, 32), 1, 1.0, 1), ((48, 24, 40), 2, 1.8, 2), ((9, 7, 16), 3, 0.7, 1),
         ((20, 12, 36), 4, 1.2, 3), ((64, 64, 64), 5, 1.5, 2), ((17, 33, 24), 6, 0.9, 1)]


@pytest.mark.parametrize("storage,outer,method", [("bf16", None, "fcg"), ("f32", None, "fcg"),
                                                  ("bf16", "f32", "fcg"),
                                                  ("bf16", None, "refine")])
def test_solver_writes_stay_inside_their_buffers(dev, storage, outer, method):
    from hfm.mg import MGBatch
    print(f"guard solve {storage} {outer} {method}", flush=True)
    alloc = GuardedAlloc(dev)
    mg = MGBatch([s[0] for s in SPECS], storage=storage, outer=outer, method=method, device=dev,
                 fuse_min_tiles=0, alloc=alloc, maxit=60)
    for i, (shape, seed, sig, nd) in enumerate(SPECS):
        sx, sy, sz, b = O.make_anisotropic_fields(shape, seed, sig, nd)
        mg.set_problem(i, sx, sy, sz, b)
    mg.setup()
    # stationary defect correction need not reach 1e-8 in 60 cycles on these
    # fields; it must stop cleanly
    ok = (1,) if method == "fcg" else (1, 2)
    info = mg.solve()
    assert np.all(np.isin(info["status"], ok))
    assert alloc.damaged() == []
    info = mg.solve(persistent=True)
    assert np.all(np.isin(info["status"], ok))
    assert alloc.damaged() == []


def test_reference_mode_kernels_write_inside_their_buffers(dev):
    """The column walkers and fills on a grid that fills no block evenly."""
    import torch
    from hfm import _lib as L, layers
    from hfm.timestep import urban_kernels
    sh = (7, 13, 70)
    rng = np.random.default_rng(3)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(dev)  # noqa: E731
    u, v, w, p, a = (T(rng.standard_normal(sh)) for _ in range(5))
    sig = T(rng.uniform(0, 2, sh))
    alloc = GuardedAlloc(dev)
    s = L.stream_handle(dev)
    k = urban_kernels(0.4, 1.0, 1 / 20)
    D, H, W = sh
    # conv27 in every form
    K27 = layers.kernel27(k["poisson"])
    for form in ("apply", "resid", "jacobi", "subtract", "axpy"):
        y = alloc.empty_like(u)
        L.call("hfm_conv27_f32", u.data_ptr(), y.data_ptr(), v.data_ptr(),
               K27.ctypes.data_as(L.C.c_void_p), float(K27[13]), D, H, W, L.FORM[form], s)
    # the fused gradient, in place on guarded copies, and the velocity fills
    outs = [alloc.empty_like(u) for _ in range(3)]
    for o, f in zip(outs, (u, v, w)):
        o.copy_(f)
    pout = alloc.empty_like(p)
    Kg = np.ascontiguousarray(np.stack([layers.kernel27(x) for x in k["adv"]]), np.float32)
    for flags in (L.GRAD_P_BC | L.GRAD_PRE_BLUFF, L.GRAD_P_BC | L.GRAD_POST_BLUFF):
        L.call("hfm_grad3_f32", p.data_ptr(), L.ptr_array([o.data_ptr() for o in outs]),
               L.ptr_array([o.data_ptr() for o in outs]), pout.data_ptr(),
               Kg.ctypes.data_as(L.C.c_void_p), sig.data_ptr(), 0.4, flags, D, H, W, s)
    scratch = alloc(2 * D * W + 2 * H * W, torch.float32)
    L.call("hfm_bc_velocity_bluff_inplace_f32", *(o.data_ptr() for o in outs), scratch.data_ptr(),
           D, H, W, 1.0, None, sig.data_ptr(), 0.4, s)
    L.call("hfm_bc_velocity_inplace_f32", *(o.data_ptr() for o in outs), scratch.data_ptr(),
           D, H, W, 1.0, None, s)
    # advection with the scalar's fill and dampings
    K4 = np.ascontiguousarray(np.stack([layers.kernel27(x) for x in (k["dif"], *k["adv"])]),
                              np.float32)
    a_out = alloc.empty_like(a)
    L.call("hfm_advect27_post_f32", L.ptr_array([u.data_ptr(), v.data_ptr(), w.data_ptr()]),
           L.ptr_array([a.data_ptr()]), L.ptr_array([a.data_ptr()]),
           L.ptr_array([a_out.data_ptr()]), 1, K4.ctypes.data_as(L.C.c_void_p), 1.0,
           sig.data_ptr(), 0.4, L.ADV_INDICATOR | (2 << L.ADV_BLUFF_SHIFT), D, H, W, s)
    v_out = [alloc.empty_like(u) for _ in range(3)]
    L.call("hfm_advect27_f32", L.ptr_array([u.data_ptr(), v.data_ptr(), w.data_ptr()]),
           L.ptr_array([u.data_ptr(), v.data_ptr(), w.data_ptr()]),
           L.ptr_array([u.data_ptr(), v.data_ptr(), w.data_ptr()]),
           L.ptr_array([o.data_ptr() for o in v_out]), 3, K4.ctypes.data_as(L.C.c_void_p), 0.5,
           D, H, W, s)
    # boundary fills, restriction and prolongations on an even grid
    for kind in L.BC.values():
        y = alloc.empty_like(u)
        L.call("hfm_bc_f32", u.data_ptr(), y.data_ptr(), D, H, W, kind, s)
    se = (8, 14, 70)
    x = T(rng.standard_normal(se))
    xc = alloc(4 * 7 * 35, torch.float32)
    L.call("hfm_restrict_mean_f32", x.data_ptr(), xc.data_ptr(), *se, 2, 2, 2, s)
    xf = alloc(int(np.prod(se)), torch.float32)
    L.call("hfm_prolong_nearest_f32", xc.data_ptr(), xf.data_ptr(), 4, 7, 35, 2, 2, 2, s)
    xu = alloc(int(np.prod(se)), torch.float32)
    L.call("hfm_prolong_upsum_f32", xc.data_ptr(), xu.data_ptr(), 4, 7, 35, 2, 2, 2, s)
    assert alloc.damaged() == []
