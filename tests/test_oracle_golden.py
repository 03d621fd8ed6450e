"""The oracle against numbers the reference's notebooks printed.

Fixtures: tests/golden/reference_outputs.json (written by
tests/golden/extract_reference_outputs.py from recorded cell outputs).
"""
import json
import os

import numpy as np
import pytest

from oracle import hfm_oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.json")))


def _sawtooth_setup(nd, n=128):
    """Initial field of ARTn/Parallel_multigrid_{1,2}D.ipynb (#c15-#c16 / #c14-#c15):
    ones on [43, 83) per axis, b = values; levels n -> 2, one sweep each."""
    shape = (n,) * nd
    T = np.zeros(shape, np.float32)
    T[(slice(43, 83),) * nd] = 1.0
    K = O.implicit_ad2d_kernel() if nd == 2 else O.implicit_ad1d_kernel()
    d = np.float32(K[(1,) * nd])
    nlev = int(np.log2(n))  # n, n/2, ..., 2
    ops = [lambda x, K=K: O.conv_same(x, K)] * nlev
    return T, K, d, nlev, ops


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sawtooth_2d_golden(dtype):
    T, K, d, nlev, ops = _sawtooth_setup(2)
    T = T.astype(dtype)
    rec = {}
    O.sawtooth_cycle(T, T, ops, [dtype(d)] * nlev, record=rec)
    w2 = rec[nlev - 1]          # 2x2 after its sweep (cell 50)
    w4 = rec[nlev - 2]          # 4x4 after its sweep (cell 52)
    np.testing.assert_array_equal(w2.ravel(), np.array(GOLD["mg2d_w2"], dtype))
    np.testing.assert_allclose(w4.ravel(), GOLD["mg2d_w4"], rtol=0, atol=5e-9)


def test_sawtooth_1d_golden():
    T, K, d, nlev, ops = _sawtooth_setup(1)
    rec = {}
    O.sawtooth_cycle(T, T, ops, [np.float32(d)] * nlev, record=rec)
    lev32 = nlev - 5          # sizes 128,64,32,... -> index of 32
    assert rec[lev32].shape == (32,)
    np.testing.assert_allclose(rec[("A_prolonged", lev32)], GOLD["mg1d_w32_cnn"], rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(rec[lev32], GOLD["mg1d_w32_jacobi"], rtol=2e-6, atol=1e-11)
    np.testing.assert_allclose(rec[nlev - 2], GOLD["mg1d_w4"], rtol=1e-6, atol=1e-8)


def test_3d_weight_layout_known_answers():
    # ARTn/Multigrid_3d_7point_source.ipynb: dt*D = (20, 10, 5) for (x, y, z)
    K = O.a7_source_kernel(dt=0.1, Dx=200.0, Dy=100.0, Dz=50.0)
    assert K.sum() == GOLD["src3d_sum"][0] == 0.0
    np.testing.assert_array_equal(K[1].ravel(), GOLD["src3d_slice"])
    assert K[1, 1, 1] == GOLD["src3d_centre"][0]
    np.testing.assert_array_equal(K.ravel(), GOLD["src3d_getweights"])
    # and the conv reads K[a,b,c] at offset (a-1, b-1, c-1): axis 0 carries Dz
    x = np.zeros((3, 3, 3)); x[0, 1, 1] = 1.0
    assert O.conv_same(x, K)[1, 1, 1] == -5.0
    x = np.zeros((3, 3, 3)); x[1, 1, 2] = 1.0
    assert O.conv_same(x, K)[1, 1, 1] == -20.0


def test_implicit_stencils_sum_to_one():
    # "Sum up the weights of filter: [1.]" (ARTn/Parallel_multigrid_2D.ipynb#c5 out, _1D#c4 out)
    assert abs(O.implicit_ad2d_kernel().sum() - 1.0) < 1e-12
    assert abs(O.implicit_ad1d_kernel().sum() - 1.0) < 1e-12
