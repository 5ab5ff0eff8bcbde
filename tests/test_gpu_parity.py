"""GPU parity: the HIP path (through the C ABI) against the oracle on the same inputs.

Tolerances (relative max-norm per conserved variable, normalised by the
variable's max magnitude):
  FP64 vs FP64 oracle: 1e-11 after up to 10 RK steps (only FMA contraction
      and summation-order differences; the exchange and the level-set fill
      are evaluated without contraction and match bit for bit).
  FP32 vs FP32 oracle: 2e-5 after up to 10 RK steps.
Exchange-only and level-set-only results are compared bit for bit.
"""
import numpy as np
import pytest

from drnum_amd import configs, hip, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-11, np.float32: 2e-5}


def tables_of(og):
    return [og.tables(p) for p in range(og.np)]


def rel_linf(gpu_states, og, field=0):
    """max over patches and variables of |gpu - oracle| / scale_v, with scale_v the
    variable's max magnitude.  For the three momenta scale_v is floored at 1% of
    sqrt(rho_max (rho E)_max), about the momentum of a flow at the speed of sound,
    so that components that vanish in exact arithmetic (transverse momentum) are
    measured against the flow's own momentum scale in any system of units.  A
    non-finite GPU value gives inf; a non-finite oracle value means the test
    problem itself is invalid and raises."""
    rho_max = e_max = 0.0
    for p in range(len(gpu_states)):
        R = og.field(p, field)
        if not np.isfinite(R).all():
            raise AssertionError("oracle state of patch %d is not finite: invalid test problem" % p)
        rho_max = max(rho_max, float(np.abs(R[0]).max()))
        e_max = max(e_max, float(np.abs(R[4]).max()))
    mom_floor = 0.01 * float(np.sqrt(rho_max * e_max))
    worst = 0.0
    for p, G in enumerate(gpu_states):
        Gd = G.astype(np.float64)
        if not np.isfinite(Gd).all():
            return float("inf")
        R = og.field(p, field).astype(np.float64)
        for v in range(5):
            scale = float(np.abs(R[v]).max())
            if 1 <= v <= 3:
                scale = max(scale, mom_floor)
            worst = max(worst, float(np.abs(Gd[v] - R[v]).max()) / scale)
    return worst


def make(w, dtype, threads=8):
    og = O.OracleGrid(w, dtype, threads=threads)
    og.init_state()
    og.compute_dependencies()
    gg = solver.GpuGrid(w, dtype)
    gg.upload_initial()
    gg.set_tables(tables_of(og))
    return og, gg


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_single_stage_sod(gpu, dtype):
    """One CartesianIterator::compute pass on a 16^3 Sod patch (no exchange)."""
    w = configs.sod(n=16)
    og, gg = make(w, dtype)
    og.copy_field(0, 1)
    gg.ctx.copy_field(0, 1)
    factor = 0.25 * w.dt
    og.compute(factor)
    gg.ctx.stage(0, 1, 2, factor)
    gg.ctx.sync()
    err = rel_linf(gg.state(2), og)
    print("single stage sod16", np.dtype(dtype).name, err, flush=True)
    assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("limiter", [configs.LIM_VANALBADA, configs.LIM_MINMOD])
def test_sod_rk_steps(gpu, dtype, limiter):
    """10 RK steps of the 16^3 Sod tube.  The unlimited second-order slopes are
    not run here: at the contact they drive the pressure negative within a few
    steps, in the oracle as on the GPU (test_smooth_shear_rk_steps covers them)."""
    w = configs.sod(n=16)
    w.limiter = limiter
    og, gg = make(w, dtype)
    for s in range(10):
        og.rk_step()
        gg.rk_step()
    err = rel_linf(gg.state(0), og)
    print("sod16 10 steps lim", limiter, np.dtype(dtype).name, err, flush=True)
    assert err <= TOL[dtype]


def smooth_shear(core=16, blocks=2):
    """Periodic double shear layer in the jet's units (rho = 1, p = 1/1.4, so
    a = 1): u = 0.5 tanh((y - 1/4)/d) below y = 1/2 and its mirror above, d two
    cells, v = 0.05 sin(2 pi x).  Mach 0.5 and no discontinuity anywhere."""
    nc = core * blocks
    h = 1.0 / nc
    d = 2.0 * h
    patches = configs.block_grid([nc] * 3, [blocks] * 3, [1.0] * 3, 2, [True] * 3, [configs.BC_SLIP] * 6)

    def init(x, y, z):
        yy = np.mod(y, 1.0)
        u = np.where(yy < 0.5, 0.5 * np.tanh((yy - 0.25) / d), -0.5 * np.tanh((yy - 0.75) / d))
        v = 0.05 * np.sin(2.0 * np.pi * x)
        p = np.full_like(x, configs.P_AMB)
        T = p / configs.gas.R
        return np.stack([p, T, u, v, np.zeros_like(x)])

    dt = 0.5 * h / 1.5
    return configs.Workload("smooth_shear%dx%d" % (blocks ** 3, core), patches, configs.RECON_UPWIND2,
                            configs.LIM_VANALBADA, [1.0] * 3, [0.25, 0.5, 1.0], dt, init)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("limiter", [configs.LIM_VANALBADA, configs.LIM_MINMOD, configs.LIM_SECONDORDER])
def test_smooth_shear_rk_steps(gpu, dtype, limiter):
    """Every limiter, the unlimited second-order slopes included, on a smooth
    flow (smooth_shear: 8 patches of 16^3 cells, exchange every stage)."""
    w = smooth_shear()
    w.limiter = limiter
    og, gg = make(w, dtype)
    for s in range(10):
        og.rk_step()
        gg.rk_step()
    err = rel_linf(gg.state(0), og)
    print("shear8x16 10 steps lim", limiter, np.dtype(dtype).name, err, flush=True)
    assert err <= TOL[dtype]


@pytest.mark.parametrize("faces", ["transmissive", "reference"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_jet_patches_rk(gpu, dtype, faces):
    """2x2x2 patches of 8^3 core cells: jet inflow, exchange every stage; the
    other domain faces transmissive, or as the reference's JetFlux has them
    (far-field on the lateral faces, a slip wall at the far end)."""
    w = configs.jet(core=8, blocks=2, faces=faces)
    og, gg = make(w, dtype)
    for s in range(10):
        og.rk_step()
        gg.rk_step()
    err = rel_linf(gg.state(0), og)
    print("jet2x8", faces, "10 steps", np.dtype(dtype).name, err, flush=True)
    assert err <= TOL[dtype]


def test_batched_equals_per_patch(gpu):
    """The batched work-list launch and the one-launch-per-patch form are bitwise identical."""
    w = configs.jet(core=8, blocks=2)
    og = O.OracleGrid(w, np.float64)
    og.compute_dependencies()
    a = solver.GpuGrid(w, np.float64)
    b = solver.GpuGrid(w, np.float64)
    for g in (a, b):
        g.upload_initial()
        g.set_tables(tables_of(og))
    for s in range(3):
        a.rk_step(hip.DRN_LAUNCH_BATCHED)
        b.rk_step(hip.DRN_LAUNCH_PER_PATCH)
    for x, y in zip(a.state(0), b.state(0)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_exchange_bitwise(gpu, dtype):
    """Patch::accessDonorDataDirect on a random state: bit-identical to the oracle."""
    w = configs.jet(core=8, blocks=2)
    og, gg = make(w, dtype)
    rng = np.random.default_rng(5)
    states = []
    for p in range(og.np):
        U = (1.0 + rng.random((5, og.cells(p)))).astype(dtype)
        og.field(p, 0)[:] = U
        states.append(U)
    gg.upload_state(states, 0)
    og.exchange(0)
    gg.ctx.exchange(0)
    for p, G in enumerate(gg.state(0)):
        assert np.array_equal(G, og.field(p, 0)), p


def test_graph_equals_eager(gpu):
    w = configs.jet(core=8, blocks=2)
    og = O.OracleGrid(w, np.float32)
    og.compute_dependencies()
    a = solver.GpuGrid(w, np.float32)
    b = solver.GpuGrid(w, np.float32)
    for g in (a, b):
        g.upload_initial()
        g.set_tables(tables_of(og))
    for s in range(5):
        a.rk_step()
    b.ctx.rk_graph_run(w.dt, w.alpha, 5)
    for x, y in zip(a.state(0), b.state(0)):
        assert np.array_equal(x, y)


def test_graph_replays_from_any_field_map(gpu):
    """A graph holds the physical buffers of the field map it was captured from.
    Calls that leave the map elsewhere (a remainder step, an eager step, a second
    call) must replay a graph of the map they start from: graph(3) twice, an
    eager step then graph(3), and graph(1) then graph(4) all equal the same
    number of eager steps bit for bit."""
    w = configs.jet(core=8, blocks=2)
    og = O.OracleGrid(w, np.float32)
    og.compute_dependencies()
    t = tables_of(og)

    def fresh():
        g = solver.GpuGrid(w, np.float32)
        g.upload_initial()
        g.set_tables(t)
        return g

    ref = fresh()
    eager = []
    for s in range(6):
        ref.rk_step()
        if s in (3, 4, 5):
            eager.append(ref.state(0))
    cases = {
        "graph3+graph3": ([("g", 3), ("g", 3)], eager[2]),
        "eager1+graph3": ([("e", 1), ("g", 3)], eager[0]),
        "graph1+graph4": ([("g", 1), ("g", 4)], eager[1]),
        "graph2+eager1+graph2": ([("g", 2), ("e", 1), ("g", 2)], eager[1]),
    }
    for name, (calls, want) in cases.items():
        print("graph case", name, flush=True)
        g = fresh()
        for kind, n in calls:
            if kind == "g":
                g.ctx.rk_graph_run(w.dt, w.alpha, n)
            else:
                g.rk_step()
        for x, y in zip(g.state(0), want):
            assert np.array_equal(x, y), name
        g.close()
    ref.close()


@pytest.mark.parametrize("graph", [False, True])
def test_upload_after_steps_is_exchanged(gpu, graph):
    """A state uploaded into a context that has already stepped starts its next
    step with the leading exchange of rungekutta.cpp:30, as in a fresh context:
    step, upload a new state, step equals upload the new state, step."""
    w = configs.jet(core=8, blocks=2)
    og = O.OracleGrid(w, np.float64)
    og.compute_dependencies()
    t = tables_of(og)
    rng = np.random.default_rng(17)
    new = [(w.initial_state(p, np.float64) * (1.0 + 0.01 * rng.random((5, P.cells)))).astype(np.float64)
           for p, P in enumerate(w.patches)]
    a = solver.GpuGrid(w, np.float64)
    a.upload_initial()
    a.set_tables(t)
    b = solver.GpuGrid(w, np.float64)
    b.set_tables(t)
    b.upload_state(new)
    if graph:
        a.ctx.rk_graph_run(w.dt, w.alpha, 2)
        a.upload_state(new)
        a.ctx.rk_graph_run(w.dt, w.alpha, 2)
        b.ctx.rk_graph_run(w.dt, w.alpha, 2)
    else:
        a.rk_step()
        a.upload_state(new)
        a.rk_step()
        b.rk_step()
    for x, y in zip(a.state(0), b.state(0)):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_core_sums(gpu, dtype):
    w = configs.jet(core=8, blocks=2)
    og, gg = make(w, dtype)
    for s in range(2):
        og.rk_step()
        gg.rk_step()
    a = gg.ctx.core_sums(0)
    b = og.core_sums(0)
    # momentum sums cancel to ~1e-6 of the mass: compare against the largest sum
    tol = (1e-12 if dtype == np.float64 else 2e-5) * np.abs(b).max()
    assert np.all(np.abs(a - b) <= tol), (a, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_every_tiling(gpu, dtype):
    """Every stage-kernel work-list choice (marching i-ranges, XCD order, cell kernel)
    on 2x2x2 patches of 16^3 core cells (planes of 324-400 cells: several chunks)."""
    w = configs.jet(core=16, blocks=2)
    og = O.OracleGrid(w, dtype, threads=8)
    og.init_state()
    og.compute_dependencies()
    for s in range(4):
        og.rk_step()
    tables = tables_of(og)
    n = solver.GpuGrid(w, dtype).ctx.num_tilings()
    for choice in range(n):
        g2 = solver.GpuGrid(w, dtype)
        g2.upload_initial()
        g2.set_tables(tables)
        g2.ctx.set_tiling(choice, hip.DRN_LAUNCH_BATCHED)
        g2.ctx.set_tiling(choice, hip.DRN_LAUNCH_PER_PATCH)
        for s in range(4):
            g2.rk_step(hip.DRN_LAUNCH_BATCHED if s % 2 == 0 else hip.DRN_LAUNCH_PER_PATCH)
        err = rel_linf(g2.state(0), og)
        print("tiling", choice, np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype], (choice, err)
        g2.close()


@pytest.mark.parametrize("arith", [0, 1])
def test_arith_modes_fp32(gpu, arith):
    """IEEE and fast arithmetic of the FP32 flux kernels against the FP32 oracle."""
    w = configs.sod(n=16)
    og, gg = make(w, np.float32)
    gg.ctx.set_arith(arith)
    for s in range(10):
        og.rk_step()
        gg.rk_step()
    err = rel_linf(gg.state(0), og)
    print("arith", arith, err, flush=True)
    assert err <= TOL[np.float32]


@pytest.mark.parametrize("arith", [hip.DRN_ARITH_IEEE, hip.DRN_ARITH_FAST])
@pytest.mark.parametrize("tiling", [11, 27, 0, 12, 22, 29, 31, 33, 38, 41])
def test_sod_keeps_mirror_symmetry(gpu, tiling, arith):
    """Sod along x between slip walls in y and z.  In the oracle the y and z
    momenta stay exactly 0 and every x plane stays uniform; the GPU must keep
    both exactly, in both FP32 arithmetic modes and for each kernel family
    (cell, z-shared, march, column, wave tile, y/z-shared, two-plane,
    persistent, i-marching z-shared).  A slip wall whose pressure rounds differently from the
    interior faces' pressure term breaks this at the wall cells."""
    n = 16
    w = configs.sod(n=n)
    _, gg = make(w, np.float32)
    gg.ctx.set_tiling(tiling)
    gg.ctx.set_arith(arith)
    for s in range(10):
        gg.rk_step()
    G = gg.state(0)[0].reshape(5, n, n, n)
    assert np.isfinite(G).all()
    assert np.count_nonzero(G[2]) == 0 and np.count_nonzero(G[3]) == 0
    for v in (0, 1, 4):
        assert np.array_equal(G[v], np.broadcast_to(G[v][:, :1, :1], G[v].shape)), v
    gg.close()


def test_autotune_keeps_state(gpu):
    w = configs.jet(core=8, blocks=2)
    og, gg = make(w, np.float32)
    before = gg.state(0)
    choice = gg.ctx.autotune(hip.DRN_LAUNCH_BATCHED)
    assert 0 <= choice < gg.ctx.num_tilings()
    choice = gg.ctx.autotune(hip.DRN_LAUNCH_PER_PATCH)
    assert 0 <= choice < gg.ctx.num_tilings()
    after = gg.state(0)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    for s in range(3):
        og.rk_step()
        gg.rk_step()
    assert rel_linf(gg.state(0), og) <= TOL[np.float32]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tiling", [11, 27, 35, 33, 0, 12, 22, 29, 31, 38])
def test_upwind1_rk_steps(gpu, dtype, tiling):
    """First-order Upwind1 (upwind1.h:29-39) through every kernel family: the Sod
    tube on one patch and the smooth shear layer over 8 patches (exchange every
    stage, fused where the kernel fuses), 10 RK steps against the oracle."""
    for w in (configs.sod(n=16), smooth_shear()):
        w.recon = configs.RECON_UPWIND1
        og, gg = make(w, dtype)
        gg.ctx.set_tiling(tiling)
        for s in range(10):
            og.rk_step()
            gg.rk_step()
        err = rel_linf(gg.state(0), og)
        print("upwind1", w.name, "tiling", tiling, np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype]


FAMILY_TILINGS = [11, 27, 35, 33, 0, 12, 22, 29, 31, 38, 41]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("limiter", [configs.LIM_MINMOD, configs.LIM_SECONDORDER])
@pytest.mark.parametrize("tiling", FAMILY_TILINGS)
def test_limiters_through_every_kernel_family(gpu, dtype, limiter, tiling):
    """minmod (minmod.h:28-31) and the unlimited second-order slopes
    (secondorder.h:28-31) through one tiling of each kernel family, 10 RK
    steps of the 8-patch smooth shear (and of the Sod tube for minmod)."""
    cases = [smooth_shear()] + ([configs.sod(n=16)] if limiter == configs.LIM_MINMOD else [])
    for w in cases:
        w.limiter = limiter
        og, gg = make(w, dtype)
        gg.ctx.set_tiling(tiling)
        for s in range(10):
            og.rk_step()
            gg.rk_step()
        err = rel_linf(gg.state(0), og)
        print("limiter", limiter, w.name, "tiling", tiling, np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype]


@pytest.mark.parametrize("tiling", FAMILY_TILINGS)
def test_fp32_ieee_arithmetic_through_every_kernel_family(gpu, tiling):
    """FP32 with IEEE division and square root (DRN_ARITH_IEEE) through one tiling
    of each kernel family (the families without an IEEE form run the cell kernel)."""
    w = smooth_shear()
    og, gg = make(w, np.float32)
    gg.ctx.set_arith(hip.DRN_ARITH_IEEE)
    gg.ctx.set_tiling(tiling)
    for s in range(10):
        og.rk_step()
        gg.rk_step()
    err = rel_linf(gg.state(0), og)
    print("ieee fp32 tiling", tiling, err, flush=True)
    assert err <= TOL[np.float32]


def z_wave(zc):
    """Two periodic patches stacked along z, 5 x 4 x zc core cells each (nk =
    zc + 4 stored cells with the seek layers), smooth in all three directions
    (Mach about 0.4).  A wavefront of the z-shared kernels holds 63 consecutive
    cells of the flattened (j,k) plane and takes the z neighbours of its cells
    from the neighbouring lanes; nk decides where the wavefront edges (lanes 0,
    1 and 63, which load the neighbours outside it) fall along the rows."""
    patches = configs.block_grid([5, 4, 2 * zc], [1, 1, 2], [1.0] * 3, 2, [True] * 3, [configs.BC_SLIP] * 6)

    def init(x, y, z):
        p = configs.P_AMB * (1.0 + 0.05 * np.sin(2.0 * np.pi * z))
        T = p / ((1.0 + 0.1 * np.cos(2.0 * np.pi * (x + z))) * configs.gas.R)
        u = 0.2 * np.sin(2.0 * np.pi * z)
        v = 0.1 * np.cos(2.0 * np.pi * x)
        w = 0.1 + 0.2 * np.sin(2.0 * np.pi * y)
        return np.stack([p, T, u, v, w])

    dt = 0.4 * (1.0 / (2 * zc)) / 1.5
    return configs.Workload("zwave%d" % zc, patches, configs.RECON_UPWIND2, configs.LIM_VANALBADA, [1.0] * 3,
                            [0.25, 0.5, 1.0], dt, init)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("zc", [3, 57, 59, 60, 61, 62, 63, 123])
def test_z_neighbours_across_wavefront_edges(gpu, dtype, zc):
    """The z-shared kernels (cell, push-chain list, persistent, i-marching) against the
    oracle, 5 RK steps, for z lengths that put the wavefront edges at every
    position relative to the rows and the patch ends."""
    w = z_wave(zc)
    for tiling in (27, 35, 33, 38, 41):
        og, gg = make(w, dtype)
        gg.ctx.set_tiling(tiling)
        for s in range(5):
            og.rk_step()
            gg.rk_step()
        err = rel_linf(gg.state(0), og)
        print("z lanes nk", zc + 4, "tiling", tiling, np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_x_shared_planes_equal_the_one_plane_kernel(gpu, dtype):
    """The z-shared cell kernel over 2 and 4 planes per work item (choices
    47-50) takes the x face i-1/2 of a cell from the wavefront of plane i-1,
    which evaluated it as its plus face from the same four cells: the states
    equal those of the one-plane kernel (choice 27) bit for bit, batched and
    per patch, with the fused z-face exchange."""
    w = configs.jet(core=16, blocks=2)
    og = O.OracleGrid(w, dtype, threads=8)
    og.compute_dependencies()
    tables = tables_of(og)

    def run(choice):
        g = solver.GpuGrid(w, dtype)
        g.upload_initial()
        g.set_tables(tables)
        g.ctx.set_tiling(choice, hip.DRN_LAUNCH_BATCHED)
        g.ctx.set_tiling(choice, hip.DRN_LAUNCH_PER_PATCH)
        for s in range(4):
            g.rk_step(hip.DRN_LAUNCH_BATCHED if s % 2 == 0 else hip.DRN_LAUNCH_PER_PATCH)
        st = g.state(0)
        g.close()
        return st

    ref = run(27)
    for choice in (47, 48, 49, 50):
        print("x-shared choice", choice, np.dtype(dtype).name, flush=True)
        for p, (a, b) in enumerate(zip(ref, run(choice))):
            assert np.array_equal(a, b), (choice, p)
