"""The exchange of the aligned z-face seek layers fused into the cell stage
kernels (drn_set_exchange_fusion): the donor's stage thread writes the value
the exchange would copy, the receiver's thread skips its store, and the
exchange after the stage runs on the other receivers.  Every stored cell must
come out bit-identical to the unfused stage + exchange, for each kernel that
fuses (cell, z-shared, persistent, i-marching) in both launch forms and both precisions, on
a block grid (jet), a periodic box (KH), the viscous operator, the level-set
fill and a nested grid, whose sides that are not aligned unit-weight copies
stay with the exchange.  Tilings of the other kernel families never fuse."""
import numpy as np
import pytest

from drnum_amd import configs, hip, host

pytestmark = pytest.mark.gpu

B, PP = hip.DRN_LAUNCH_BATCHED, hip.DRN_LAUNCH_PER_PATCH

CASES = {
    # 8 patches; each has one z side to an interior neighbour: 8 x 8 core columns x 2 layers
    "jet": (lambda: configs.jet(core=8, blocks=2), 8 * 8 * 8 * 2),
    # 27 periodic patches, both z sides of every patch: 8 x 8 x 2 x 2
    "kh": (lambda: configs.kelvin_helmholtz(core=8, blocks=3), 27 * 8 * 8 * 2 * 2),
    "shear": (lambda: configs.viscous_shear(core=8, blocks=(2, 2, 2)), 8 * 8 * 8 * 2 * 2),
    # one periodic patch, its own z neighbour: it pushes into its own seek layers
    "kh1": (lambda: configs.kelvin_helmholtz(core=8, blocks=1), 1 * 8 * 8 * 2 * 2),
    # 2-cell cores: the same donor cells push to both z neighbours (one patch, periodic)
    "kh_thin": (lambda: configs.kelvin_helmholtz(core=2, blocks=2), 8 * 2 * 2 * 2 * 2),
    # 12 patches of 40^3 core (3 x 2 x 2): the z sides between the two layers
    "spheres": (lambda: configs.levelset_spheres(core=40), 38400),
    # overlapping and nested patches of 12-24 cells (sizes drawn per patch): the z sides
    # that are aligned unit-weight copies fuse, the others stay with the exchange
    "uneven": (lambda: configs.uneven_batch(n_patches=30, size_range=(12, 24)), None),
}
# (tiling, launch form) of each fusing kernel: cell, z-shared, persistent,
# z-shared on the push-chain list, per-patch z-shared, the i-marching z-shared
# kernel (16-plane items; capped at 8 waves per SIMD; whole patches per patch),
# the z-shared kernel over 4 and 2 planes per item sharing x faces (push-chain
# list; per patch)
FORMS = [(11, B), (27, B), (33, B), (35, B), (28, PP), (38, B), (41, B), (36, PP), (47, B), (49, PP)]


def setup(w, dtype, fuse):
    hg = host.HostGrid.from_workload(w, fp64_geometry=True)
    hg.compute_dependencies(w.periodic)
    ctx = hip.Context(dtype)
    ctx.set_operator(w.recon, w.limiter, w.viscous)
    hg.to_device(ctx)
    hg.upload_exchange(ctx)
    if w.spheres is not None:
        hg.levelset_spheres(w.spheres, w.ls_inner_layers, 0, w.ls_min_inner_rel)
        hg.upload_levelset(ctx)
    for p in range(len(w.patches)):
        ctx.upload(p, 0, w.initial_state(p, dtype))
    ctx.set_exchange_fusion(fuse)
    return hg, ctx


def run(w, dtype, fuse, tiling, launch, steps=3, graph=False):
    _, ctx = setup(w, dtype, fuse)
    ctx.set_tiling(tiling, launch)
    if graph:
        ctx.rk_graph_run(w.dt, w.alpha, steps)
    else:
        for _ in range(steps):
            ctx.rk_step(w.dt, w.alpha, launch)
    out = [ctx.download(p, 0) for p in range(len(w.patches))]
    n = ctx.exchange_fused_cells(launch)
    ctx.close()
    return out, n


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", list(CASES))
def test_fused_exchange_is_bitwise_the_exchange(gpu, case, dtype):
    make, expect_fused = CASES[case]
    w = make()
    for tiling, launch in FORMS:
        print("case", case, np.dtype(dtype).name, "tiling", tiling, "launch", launch, flush=True)
        ref, n_ref = run(w, dtype, False, tiling, launch)
        got, n = run(w, dtype, True, tiling, launch)
        assert n_ref == 0
        if expect_fused is not None:
            assert n == expect_fused
        for p, (a, b) in enumerate(zip(ref, got)):
            assert np.isfinite(a).all()
            assert np.array_equal(a, b), (case, tiling, launch, p, float(np.abs(a - b).max()))


def test_fused_exchange_in_the_graph(gpu):
    w = configs.jet(core=8, blocks=2)
    ref, _ = run(w, np.float32, False, 27, B, steps=4)
    got, n = run(w, np.float32, True, 27, B, steps=4, graph=True)
    assert n == 8 * 8 * 8 * 2
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_other_kernel_families_do_not_fuse_and_agree(gpu):
    """a march tiling with fusion on takes the full exchange: the same states as
    with fusion off"""
    w = configs.jet(core=8, blocks=2)
    ref, _ = run(w, np.float32, False, 0, B)
    got, n = run(w, np.float32, True, 0, B)
    assert n == 0
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_stage_alone_never_fuses(gpu):
    """drn_stage leaves every receiver to drn_exchange: a stage followed by the
    exchange equals the same with fusion off, seek cells included"""
    w = configs.jet(core=8, blocks=2)
    outs = []
    for fuse in (False, True):
        _, ctx = setup(w, np.float64, fuse)
        ctx.set_tiling(27)
        ctx.exchange(0)
        ctx.copy_field(0, 1)
        ctx.stage(0, 1, 2, 0.25 * w.dt)
        outs.append([ctx.download(p, 2) for p in range(len(w.patches))])
        ctx.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
