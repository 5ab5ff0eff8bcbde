"""Turned (non-axis-aligned) patches on the GPU, through the C ABI.

The device exchange turns the momentum of each turned donor's weighted sum
with the donor's matrix (drn_exchange_set_turn; Patch::accessDonorDataDirect,
patch.cpp:617-645) in the multi-donor gather kernel.

  * the exchange equals the oracle's bit for bit on random states: the star
    of 90-degree turned arms (exact signed permutations, unit weights) and the
    star with a 10-degree tilted arm (trilinear weights, general matrices,
    and relations between two turned patches), in FP64 and FP32;
  * RK steps of a density pulse crossing the turned interfaces match the
    oracle within the parity tolerances, with the tables from the oracle and
    from the product's host builder (libdrnum.so, drnh_grid_upload_exchange);
  * uniform flow stays uniform across the turned interfaces on the device.
"""
import numpy as np
import pytest

from drnum_amd import configs, hip, host, solver
from oracle import oracle as O

from test_gpu_parity import TOL, make, rel_linf, tables_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tilt", [0.0, 10.0])
def test_turned_exchange_bitwise(gpu, dtype, tilt):
    w = configs.rotated_star(tilt_deg=tilt)
    og = O.OracleGrid(w, dtype)
    og.compute_dependencies()
    gg = solver.GpuGrid(w, dtype)
    gg.set_tables(tables_of(og))
    rng = np.random.default_rng(7)
    for p in range(len(w.patches)):
        U = og.field(p, 0)
        U[:] = (1.0 + 0.5 * rng.standard_normal(U.shape)).astype(dtype)
        gg.ctx.upload(p, 0, U.copy())
    print("turned exchange tilt", tilt, np.dtype(dtype).name, flush=True)
    og.exchange(0)
    gg.ctx.exchange(0)
    for p in range(len(w.patches)):
        np.testing.assert_array_equal(gg.ctx.download(p, 0), og.field(p, 0))
    gg.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("launch", [hip.DRN_LAUNCH_BATCHED, hip.DRN_LAUNCH_PER_PATCH])
def test_pulse_crosses_turned_interfaces(gpu, dtype, launch):
    w = configs.rotated_star(steps=10)
    og, gg = make(w, dtype)
    for _ in range(w.steps):
        og.rk_step()
        gg.rk_step(launch)
    err = rel_linf(gg.state(0), og)
    print("rotated star 10 steps", np.dtype(dtype).name, launch, err, flush=True)
    gg.close()
    assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_host_tables_with_turns_drive_the_device(gpu, dtype):
    w = configs.rotated_star(tilt_deg=10.0, arms="x", steps=5)
    og = O.OracleGrid(w, dtype, threads=8)
    og.init_state()
    og.compute_dependencies()
    hg = host.HostGrid.from_workload(w, fp64_geometry=(dtype == np.float64))
    hg.compute_dependencies(w.periodic)
    ctx = hip.Context(dtype)
    ctx.set_operator(w.recon, w.limiter)
    hg.to_device(ctx)
    hg.upload_exchange(ctx)
    for p in range(len(w.patches)):
        ctx.upload(p, 0, w.initial_state(p, dtype))
    for _ in range(w.steps):
        og.rk_step()
        ctx.rk_step(w.dt, w.alpha)
    err = rel_linf([ctx.download(p, 0) for p in range(len(w.patches))], og)
    print("host-table path, tilted arm, 5 steps", np.dtype(dtype).name, err, flush=True)
    ctx.close()
    assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_uniform_flow_stays_uniform_on_the_device(gpu, dtype):
    w = configs.rotated_star(pulse=0.0, steps=3)
    og, gg = make(w, dtype)
    U0 = gg.state(0)
    for _ in range(w.steps):
        gg.rk_step()
    eps = np.finfo(dtype).eps
    for p, U in enumerate(gg.state(0)):
        np.testing.assert_allclose(U, U0[p], rtol=64 * eps, atol=64 * eps * np.abs(U0[p]).max())
    gg.close()
