"""Patches with fewer than 3 cells along an axis: the reference skips that
axis' faces and walls (cartesianiterator.h:147-308, `if (m_NumI > 2)` etc.), the
y cell gradient of the viscous flux is 0 there (cartesianpatch_common.h:230) and
the x and z gradients become one-sided.  Every kernel family must take the same
branches: GPU against the oracle on single patches and a two-patch pair with
1- and 2-cell axes, for a choice of every kind of stage kernel."""
import copy

import numpy as np
import pytest

from drnum_amd import configs, hip
from test_gpu_parity import TOL, make, rel_linf

pytestmark = pytest.mark.gpu

SHAPES = [((24, 2, 5), (1, 1, 1)), ((20, 1, 1), (1, 1, 1)), ((3, 2, 24), (1, 1, 1)), ((2, 24, 2), (1, 1, 1)),
          ((5, 6, 1), (1, 1, 1)), ((24, 2, 6), (2, 1, 1))]
# one choice of each kernel family: cell, z-shared, march, column, wave tile, y/z-shared, two-plane, persistent
TILINGS = [11, 27, 0, 12, 22, 29, 31, 33]


def thin(cells, blocks, h=1e-3, viscous=False):
    L = [h * n for n in cells]
    patches = configs.block_grid(list(cells), list(blocks), L, 2, [False] * 3, [configs.BC_SLIP] * 6)

    def init(x, y, z):
        s = np.sin(2 * np.pi * x / L[0]) + 0.5 * np.cos(2 * np.pi * (y / L[1] + z / L[2]))
        p = 1.0e5 * (1.0 + 0.05 * s)
        T = 300.0 * (1.0 + 0.02 * np.cos(2 * np.pi * x / L[0]))
        return np.stack([p, T, 20.0 * s, 10.0 * np.cos(2 * np.pi * x / L[0]), 5.0 * s])

    dt = 0.3 * h / (40.0 + np.sqrt(1.4 * 287.2 * 300.0))
    return configs.Workload("thin", patches, configs.RECON_UPWIND2, configs.LIM_VANALBADA, [0.0] * 3,
                            [0.25, 0.5, 1.0], dt, init, viscous=viscous)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("cells,blocks", SHAPES)
def test_thin_patches_every_kernel_family(gpu, dtype, cells, blocks):
    w = thin(cells, blocks)
    og, _ = make(w, dtype)
    for s in range(4):
        og.rk_step()
    for choice in TILINGS:
        for launch in (hip.DRN_LAUNCH_BATCHED, hip.DRN_LAUNCH_PER_PATCH):
            print("thin", cells, blocks, np.dtype(dtype).name, "tiling", choice, "launch", launch, flush=True)
            _, gg = make(w, dtype)
            gg.ctx.set_tiling(choice, launch)
            for s in range(4):
                gg.rk_step(launch=launch)
            err = rel_linf(gg.state(0), og)
            assert err <= TOL[dtype], (choice, launch, err)


@pytest.mark.parametrize("cells,blocks", SHAPES)
def test_thin_patches_viscous(gpu, cells, blocks):
    w = thin(cells, blocks, h=1e-6, viscous=True)
    og, gg = make(w, np.float64)
    wi = copy.copy(w)
    wi.viscous = False
    oi, _ = make(wi, np.float64)
    for s in range(4):
        og.rk_step()
        oi.rk_step()
        gg.rk_step()
    print("thin viscous", cells, blocks, rel_linf(gg.state(0), og), flush=True)
    assert rel_linf(gg.state(0), og) <= TOL[np.float64]
    # the viscous terms act on these shapes
    assert max(float(np.abs(og.field(p, 0) - oi.field(p, 0)).max()) for p in range(og.np)) > 0
