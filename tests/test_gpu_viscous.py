"""GPU parity of the viscous flux (CompressibleViscFlux added to every interior
face flux after AUSM+, the EaFlux composition of drnumCylinder/main.h:127-163)
against the oracle, on the SI-unit shear layer where viscosity matters.
Tolerances as the inviscid parity tests: FP64 1e-11, FP32 2e-5 (relative max
norm per conserved variable)."""
import copy

import numpy as np
import pytest

from drnum_amd import configs, hip
from test_gpu_parity import TOL, make, rel_linf

pytestmark = pytest.mark.gpu


def shear(**kw):
    return configs.viscous_shear(core=8, blocks=(2, 2, 2), **kw)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_viscous_stage_matches_oracle(gpu, dtype):
    w = shear()
    og, gg = make(w, dtype)
    og.copy_field(0, 1)
    gg.ctx.copy_field(0, 1)
    factor = 0.25 * w.dt
    og.compute(factor)
    gg.ctx.stage(0, 1, 2, factor)
    gg.ctx.sync()
    err = rel_linf(gg.state(2), og)
    print("viscous stage", np.dtype(dtype).name, err, flush=True)
    assert err <= TOL[dtype]
    # the viscous part of one quarter stage is ~1e-4 of the state: far above the
    # FP64 tolerance (the FP32 check of its size is in the RK test below)
    wi = copy.copy(w)
    wi.viscous = False
    _, gi = make(wi, dtype)
    gi.ctx.copy_field(0, 1)
    gi.ctx.stage(0, 1, 2, factor)
    gi.ctx.sync()
    assert rel_linf(gi.state(2), og) > 1e4 * TOL[np.float64]


@pytest.mark.parametrize("dtype,arith", [(np.float64, hip.DRN_ARITH_IEEE), (np.float32, hip.DRN_ARITH_FAST),
                                         (np.float32, hip.DRN_ARITH_IEEE)])
@pytest.mark.parametrize("launch", [hip.DRN_LAUNCH_BATCHED, hip.DRN_LAUNCH_PER_PATCH])
def test_viscous_rk_steps(gpu, dtype, arith, launch):
    w = shear()
    og, gg = make(w, dtype)
    gg.ctx.set_arith(arith)
    wi = copy.copy(w)
    wi.viscous = False
    _, gi = make(wi, dtype)
    gi.ctx.set_arith(arith)
    for s in range(5):
        og.rk_step()
        gg.rk_step(launch=launch)
        gi.rk_step(launch=launch)
    err = rel_linf(gg.state(0), og)
    effect = rel_linf(gi.state(0), og)
    print("viscous rk", np.dtype(dtype).name, arith, launch, err, "viscous effect", effect, flush=True)
    assert err <= TOL[dtype]
    assert effect > 10 * TOL[dtype]


def test_viscous_tilings_agree(gpu):
    """With the viscous flux every tiling choice runs the z-shared kernel on its
    list order: any two choices give the same bits."""
    w = shear()
    ref = None
    for choice in (11, 20, 22, 27, 28, 33):
        og, gg = make(w, np.float32)
        gg.ctx.set_tiling(choice, hip.DRN_LAUNCH_BATCHED)
        for s in range(2):
            gg.rk_step()
        st = gg.state(0)
        if ref is None:
            ref = st
        for a, b in zip(ref, st):
            assert np.array_equal(a, b), choice
