"""Pins the oracle's per-cell flags (CPU only): inactive cells
(Patch::deactivate, patch_common.h:505; the update skips them,
cartesianiterator.h:315) and inside cells, whose boundary faces are split
faces (Patch::isSplitFace, patch_common.h:500-503) where Upwind2 keeps first
order (upwind2.h:41).

  * flags that put every cell on the same side change no bit of the residual;
  * cells alternating inside/outside along one axis make every face normal to
    it a split face: on a state that varies along that axis only, the Upwind2
    residual then equals the Upwind1 residual bit for bit (a zero slope adds
    nothing), while without the flags it differs;
  * a cell three or more cells away from every inside boundary keeps its
    unflagged residual bit for bit;
  * inactive cells keep their state bit for bit through RK steps while the
    active cells move.
"""
import numpy as np
import pytest

from drnum_amd import configs, gas
from oracle import oracle as O

from refsol import ijk

TRANS = configs.BC_TRANSMISSIVE


def one_patch(n, recon=configs.RECON_UPWIND2, axis=0, dt=2e-4):
    P = configs.Patch(n[0], n[1], n[2], [0] * 6, [0.0, 0.0, 0.0], [0.1 * x for x in n], [TRANS] * 6,
                      np.zeros((6, configs.NPAR)), [0, 0, 0])

    def init(x, y, z):
        s = (x, y, z)[axis]
        rho = 1.0 + 0.3 * np.sin(7.0 * s)
        p = 1.0 + 0.2 * np.cos(5.0 * s)
        u = 0.3 * np.sin(3.0 * s)
        zero = np.zeros_like(x)
        vel = [zero, zero, zero]
        vel[axis] = u
        return np.stack([p, p / (rho * gas.R), vel[0], vel[1], vel[2]])
    return configs.Workload("t", [P], recon, configs.LIM_VANALBADA, [0.0] * 3, [0.25, 0.5, 1.0], dt, init)


def residual(w, dtype, flags=None):
    g = O.OracleGrid(w, dtype)
    g.init_state()
    if flags is not None:
        g.set_cell_flags(0, flags)
    return g.residual(0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_flags_on_one_side_change_nothing(dtype):
    w = one_patch((9, 8, 7))
    n = w.patches[0].cells
    ref = residual(w, dtype)
    for flag in (0, 2):
        np.testing.assert_array_equal(residual(w, dtype, np.full(n, flag, np.uint8)), ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_alternating_inside_cells_make_the_reconstruction_first_order(dtype, axis):
    w2 = one_patch((10, 9, 8), configs.RECON_UPWIND2, axis)
    w1 = one_patch((10, 9, 8), configs.RECON_UPWIND1, axis)
    P = w2.patches[0]
    idx = ijk(P, np.arange(P.cells))[axis]
    flags = np.where(idx % 2 == 0, 2, 0).astype(np.uint8)
    got = residual(w2, dtype, flags)
    np.testing.assert_array_equal(got, residual(w1, dtype))
    assert not np.array_equal(residual(w2, dtype), got)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cells_away_from_the_inside_boundary_keep_their_residual(dtype):
    """A cell's residual reads the flags of the cells up to two away along each
    axis (the outer cells i0 of its faces' reconstructions); where all of them
    equal its own, nothing changes."""
    w = one_patch((12, 11, 10), axis=1)
    P = w.patches[0]
    i, j, k = ijk(P, np.arange(P.cells))
    inside = (i >= 4) & (i < 8) & (j >= 3) & (j < 7) & (k >= 2) & (k < 6)
    F = inside.reshape(P.ni, P.nj, P.nk)
    same = np.ones_like(F)
    for ax in range(3):
        n = F.shape[ax]
        for s in (-2, -1, 1, 2):
            src = np.take(F, np.clip(np.arange(n) + s, 0, n - 1), axis=ax)
            same &= src == F
    mask = same.ravel()
    ref = residual(w, dtype)
    got = residual(w, dtype, np.where(inside, 2, 0).astype(np.uint8))
    assert mask.sum() > 100 and (~mask).sum() > 100
    np.testing.assert_array_equal(got[:, mask], ref[:, mask])
    assert not np.array_equal(got[:, ~mask], ref[:, ~mask])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_inactive_cells_keep_their_state(dtype):
    w = one_patch((9, 8, 7), axis=2)
    P = w.patches[0]
    rng = np.random.default_rng(4)
    inactive = rng.random(P.cells) < 0.15
    g = O.OracleGrid(w, dtype)
    g.init_state()
    g.set_cell_flags(0, np.where(inactive, 1, 0).astype(np.uint8))
    U0 = g.field(0, 0).copy()
    for _ in range(3):
        g.rk_step()
    U = g.field(0, 0)
    np.testing.assert_array_equal(U[:, inactive], U0[:, inactive])
    assert np.all(np.any(U[:, ~inactive] != U0[:, ~inactive], axis=0).sum() > 0.9 * (~inactive).sum())
