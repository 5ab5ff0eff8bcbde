"""The oracle's Patch::computeVariableDifference (patch.cpp:733-746): the
square root of the largest squared difference and of the sum of squared
differences between two variables, over every cell of a patch, in the state
type.  Pinned by a constant offset (exact) and by numpy on random fields."""
import numpy as np
import pytest

from drnum_amd import configs
from oracle import oracle as O


def grid(dtype):
    w = configs.jet(core=6, blocks=2)
    og = O.OracleGrid(w, dtype)
    og.init_state()
    return w, og


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_constant_offset(dtype):
    w, og = grid(dtype)
    N = w.patches[3].cells
    og.field(3, 0)[0, :] = 0
    og.field(3, 1)[0, :] = dtype(0.375)
    mx, l2 = og.variable_difference(3, 0, 0, 1, 0)
    assert mx == 0.375
    assert l2 == pytest.approx(0.375 * np.sqrt(N), rel=1e-5 if dtype == np.float32 else 1e-13)
    # the same variable of the same field: zero
    assert og.variable_difference(3, 1, 0, 1, 0) == (0.0, 0.0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_fields_against_numpy(dtype):
    w, og = grid(dtype)
    rng = np.random.default_rng(3)
    for p in (0, 5):
        a = og.field(p, 0)
        b = og.field(p, 2)
        a[:] = rng.standard_normal(a.shape).astype(dtype)
        b[:] = rng.standard_normal(b.shape).astype(dtype)
        for v1, v2 in ((0, 0), (1, 4)):
            d = (b[v2] - a[v1]).astype(dtype)
            sq = (d * d).astype(dtype)
            mx, l2 = og.variable_difference(p, 0, v1, 2, v2)
            assert mx == float(np.sqrt(sq.max()).astype(dtype))
            assert l2 == pytest.approx(float(np.sqrt(sq.astype(np.float64).sum())),
                                       rel=1e-4 if dtype == np.float32 else 1e-12)
