"""The no-allocation velocity fill against the workspace form, through the C ABI (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(dev, shape, seed):
    import torch
    rng = np.random.default_rng(seed)
    return [torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(dev)
            for _ in range(3)]


def _fill(dev, fields, shape, inflow, noalloc):
    import torch
    from hfm import _lib as L
    u, v, w = (f.clone() for f in fields)
    D, H, W = shape
    s = L.stream_handle(dev)
    ip = 0 if inflow is None else inflow.data_ptr()
    if noalloc:
        L.call("hfm_bc_velocity_inplace_noalloc_f32", u.data_ptr(), v.data_ptr(), w.data_ptr(),
               D, H, W, 1.25, ip, s)
    else:
        scratch = torch.full((2 * D * W + 2 * H * W,), float("nan"), device=dev)
        L.call("hfm_bc_velocity_inplace_f32", u.data_ptr(), v.data_ptr(), w.data_ptr(),
               scratch.data_ptr(), D, H, W, 1.25, ip, s)
    return u, v, w


def test_noalloc_fill_matches_workspace_fill_twice(dev):
    """Two calls in a row, the second on a smaller grid: the shared scratch keeps the
    first call's rows, which the second must not read."""
    import torch
    for shape, seed, with_inflow in (((24, 40, 72), 1, False), ((6, 9, 17), 2, True)):
        fields = _fields(dev, shape, seed)
        inflow = (torch.linspace(0.1, 2.0, shape[0], device=dev) if with_inflow else None)
        ref = _fill(dev, fields, shape, inflow, noalloc=False)
        got = _fill(dev, fields, shape, inflow, noalloc=True)
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.equal(a, b)


def test_noalloc_fill_rejects_grids_beyond_its_scratch(dev):
    import torch
    from hfm import _lib as L
    u = torch.zeros(1, device=dev)
    D, H, W = 4, 1024, 1024        # 2 D W + 2 H W = 2105344 > 2^20 floats
    assert 2 * D * W + 2 * H * W > L.BC_NOALLOC_SCRATCH_FLOATS
    with pytest.raises(L.HfmError, match="scratch floats"):
        L.call("hfm_bc_velocity_inplace_noalloc_f32", u.data_ptr(), u.data_ptr(), u.data_ptr(),
               D, H, W, 1.0, 0, L.stream_handle(dev))
