"""drn_variable_difference, the device form of Patch::computeVariableDifference
(patch.cpp:733-746): max_norm equals the oracle's bit for bit; l2_norm, whose
sum the device accumulates in FP64, equals numpy's FP64 sum of the same
state-type squares and the oracle's state-type sum to its rounding."""
import numpy as np
import pytest

from drnum_amd import configs, hip, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_variable_difference_matches_the_oracle(gpu, dtype):
    from oracle import oracle as O

    w = configs.jet(core=8, blocks=2)
    og = O.OracleGrid(w, dtype)
    og.init_state()
    gg = solver.GpuGrid(w, dtype)
    rng = np.random.default_rng(11)
    for p in range(len(w.patches)):
        for f in (0, 2):
            U = og.field(p, f)
            U[:] = (1.0 + 0.3 * rng.standard_normal(U.shape)).astype(dtype)
            gg.ctx.upload(p, f, U.copy())
    for p in (0, 3, 7):
        for v1, v2 in ((0, 0), (2, 4)):
            print("patch", p, "vars", v1, v2, np.dtype(dtype).name, flush=True)
            mx, l2 = gg.ctx.variable_difference(p, 0, v1, 2, v2)
            omx, ol2 = og.variable_difference(p, 0, v1, 2, v2)
            assert mx == omx
            d = (og.field(p, 2)[v2] - og.field(p, 0)[v1]).astype(dtype)
            sq = (d * d).astype(dtype).astype(np.float64)
            assert l2 == pytest.approx(float(np.sqrt(sq.sum())), rel=1e-14)
            assert l2 == pytest.approx(ol2, rel=1e-4 if dtype == np.float32 else 1e-12)
    with pytest.raises(RuntimeError, match="variable index"):
        gg.ctx.variable_difference(0, 0, 5, 2, 0)
    gg.close()


def cfl_reference_x(U, dt, dx, dtype):
    """The jet demo's CFL of every cell (applications/drnumJetDemo/main.h:196-205)
    evaluated in the state type as the reference evaluates it:
    PerfectGas::conservativeToPrimitive (perfectgas.h:84-92, whose 1.0/rho,
    0.5*|u|^2 and /cv are double expressions), a = sqrt(gamma R T), then
    max(|u|*dt/dx, |u+a|*dt/dx, |u-a|*dt/dx) with the x velocity only."""
    f = np.dtype(dtype).type
    r, m1, m2, m3, e = (U[v].astype(dtype) for v in range(5))
    ir = (1.0 / r.astype(np.float64)).astype(dtype)
    u, v, w = m1 * ir, m2 * ir, m3 * ir
    ke = (u * u + v * v + w * w).astype(dtype)
    T = ((((e * ir).astype(np.float64)) - 0.5 * ke.astype(np.float64)) / 717.5).astype(dtype)
    a = np.sqrt(f(1.4) * f(287.2) * T).astype(dtype)
    dt, dx = f(dt), f(dx)
    c = np.maximum(np.abs(u) * dt / dx, np.abs(u + a) * dt / dx)
    return np.maximum(c, np.abs(u - a) * dt / dx)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cfl_is_the_jet_demos_x_only_number(gpu, dtype):
    """drn_diagnostics' CFL is the reference's x-only number and drn_cfl_axes the
    three-axis extension.  The flow moves mostly along y and z, so the two
    differ by far more than rounding; each equals its own numpy evaluation."""
    w = configs.jet(core=8, blocks=2)
    gg = solver.GpuGrid(w, dtype)
    rng = np.random.default_rng(5)
    dt = 1.3e-3
    best_x, best_3 = 0.0, 0.0
    for p, P in enumerate(w.patches):
        n = P.cells
        rho = 1.0 + 0.2 * rng.random(n)
        vel = np.stack([0.05 * rng.standard_normal(n), 0.8 + 0.1 * rng.standard_normal(n),
                        -0.6 + 0.1 * rng.standard_normal(n)])
        T = (1.0 / 1.4 / 287.2) * (1.0 + 0.1 * rng.random(n))
        U = np.empty((5, n))
        U[0] = rho
        U[1:4] = rho * vel
        U[4] = rho * (717.5 * T + 0.5 * (vel * vel).sum(axis=0))
        U = U.astype(dtype)
        gg.ctx.upload(p, 0, U.copy())
        dx, dy, dz = solver.patch_spacing(P, dtype)
        best_x = max(best_x, float(cfl_reference_x(U, dt, dx, dtype).max()))
        Ud = U.astype(np.float64)
        ud = Ud[1:4] / Ud[0]
        Td = (Ud[4] / Ud[0] - 0.5 * (ud * ud).sum(axis=0)) / 717.5
        ad = np.sqrt(1.4 * 287.2 * Td)
        d = np.array([dx, dy, dz])[:, None]
        best_3 = max(best_3, float(((np.abs(ud) + ad) * dt / d).max()))
    print("cfl", np.dtype(dtype).name, best_x, best_3, flush=True)
    diag = gg.ctx.diagnostics(0, dt)
    assert diag["cfl"] == pytest.approx(best_x, rel=2e-7 if dtype == np.float32 else 1e-15)
    assert gg.ctx.cfl_axes(0, dt) == pytest.approx(best_3, rel=1e-14)
    assert best_3 > 1.3 * best_x
    gg.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nonfinite_values_are_counted(gpu, dtype):
    """drn_diagnostics counts every NaN and infinity of the state over all stored
    cells (seek layers included), so that a run can refuse a diverged result;
    a clean state counts none."""
    w = configs.jet(core=8, blocks=2)
    gg = solver.GpuGrid(w, dtype)
    gg.upload_initial()
    assert gg.ctx.diagnostics(0, w.dt)["nonfinite"] == 0
    states = gg.state(0)
    n_last = w.patches[7].cells - 1
    bad = [(0, 0, 5, np.nan), (3, 4, 17, np.inf), (7, 2, n_last, -np.inf), (5, 1, 0, np.nan), (5, 3, 0, np.nan)]
    for p, v, cell, x in bad:
        states[p][v, cell] = x
    gg.upload_state(states, 0)
    d = gg.ctx.diagnostics(0, w.dt)
    print("non-finite values counted", d["nonfinite"], np.dtype(dtype).name, flush=True)
    assert d["nonfinite"] == len(bad)
    gg.close()
