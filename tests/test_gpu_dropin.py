"""The C++ drop-in (drnum_amd/host/iterators.h) driven the way a DrNUM
application drives the reference's GPU iterator: drnum_run reads a grid file
in the reference's format, builds HIP_CartesianIterator<5, AusmPlus<5,
Upwind2<5, VanAlbada>, PerfectGas>> and steps it with RungeKutta (the
reference's loop over the virtual iterator interface), HIP_RungeKutta or a
captured graph.

  * every mode gives the oracle's result (FP64 1e-11, FP32 2e-5 relative max
    norm) on the jet and on spheres with the level-set post operation;
  * the three modes give bit-identical states (the same kernels on the same
    inputs; only the number of copies and exchanges differs);
  * per-cell flags set on the host patches (deactivate, setSplitFaces) reach
    the device and give the oracle's result with the same flags.
"""
import os
import subprocess

import numpy as np
import pytest

from drnum_amd import configs
from oracle import oracle as O

from casefiles import case_file_bytes, grid_file_text
from test_gpu_parity import TOL, rel_linf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "drnum_amd", "lib", "drnum_run")

pytestmark = pytest.mark.gpu


def run_app(tmp_path, w, dtype, steps, mode, levelset=None, extra=(), tag="", stdout=None):
    grid = tmp_path / "grid"
    case = tmp_path / "case.bin"
    out = tmp_path / ("out_%s%s.bin" % (mode, tag))
    grid.write_text(grid_file_text(w))
    case.write_bytes(case_file_bytes(w, dtype, levelset))
    args = [APP, str(grid), str(case), str(out), str(steps), repr(float(w.dt)), mode]
    if np.dtype(dtype) == np.float64:
        args.append("f64")
    if levelset is not None:
        args.append("levelset")
    args += list(extra)
    print("drnum_run", mode, np.dtype(dtype).name, *extra, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    if stdout is not None:
        stdout.append(r.stdout)
    raw = np.frombuffer(out.read_bytes(), dtype=dtype)
    states, o = [], 0
    for P in w.patches:
        n = 5 * P.cells
        states.append(raw[o:o + n].reshape(5, P.cells))
        o += n
    assert o == raw.size
    return states


def oracle_after(w, dtype, steps, levelset=None, ls_op=0):
    og = O.OracleGrid(w, dtype, threads=8)
    og.init_state()
    og.compute_dependencies()
    if levelset is not None:
        og.levelset(levelset[0], n_inner=levelset[1], n_outer=0, min_inner_rel=levelset[2])
        og.set_levelset_op(ls_op)
    for _ in range(steps):
        og.rk_step()
    return og


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_jet_all_modes(gpu, tmp_path, dtype):
    w = configs.jet(core=8, blocks=2)
    og = oracle_after(w, dtype, 6)
    res = {m: run_app(tmp_path, w, dtype, 6, m) for m in ("rk", "hip", "graph")}
    for m, st in res.items():
        err = rel_linf(st, og)
        print("drop-in jet", m, np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype]
    for m in ("hip", "graph"):
        for a, b in zip(res["rk"], res[m]):
            assert np.array_equal(a, b), m


@pytest.mark.parametrize("swall", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_spheres_levelset(gpu, tmp_path, dtype, swall):
    """HIP_LevelSetBC<5, OP, real> (the slip wall) and
    HIP_LevelSetBC<5, OP, real, LSOBCCompressibleSWallOp> (the simple wall)"""
    w = configs.levelset_spheres(core=16, blocks=(2, 1, 1), n_spheres=2, r_range=(2.5, 3.5))
    ls = (w.spheres, 3, 0.1)
    og = oracle_after(w, dtype, 5, ls, ls_op=1 if swall else 0)
    extra = ("swall",) if swall else ()
    res = {m: run_app(tmp_path, w, dtype, 5, m, ls, extra=extra, tag="sw" if swall else "") for m in ("rk", "hip")}
    for m, st in res.items():
        err = rel_linf(st, og)
        print("drop-in spheres", m, "swall" if swall else "slip", np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype]
    for a, b in zip(res["rk"], res["hip"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_restart_file_continues_bitwise(gpu, tmp_path, dtype):
    """3 steps written as a DrNUM restart file (PatchGrid::writeData), then 2 steps
    started from it (PatchGrid::readData) give the 5-step state bit for bit."""
    from drnum_amd import host

    w = configs.jet(core=8, blocks=2)
    f64 = ["f64"] if np.dtype(dtype) == np.float64 else []
    grid = tmp_path / "grid"
    case = tmp_path / "case.bin"
    grid.write_text(grid_file_text(w))
    case.write_bytes(case_file_bytes(w, dtype))

    def app(steps, out, *extra):
        r = subprocess.run([APP, str(grid), str(case), str(out), str(steps), repr(float(w.dt)), "hip"] + f64 +
                           list(extra), capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        return np.frombuffer(out.read_bytes(), dtype=dtype)

    first = app(3, tmp_path / "a.bin", "dnd=%s" % (tmp_path / "rs"))
    shapes = [(5, P.cells) for P in w.patches]
    fields, t, prec = host.dnd_read(str(tmp_path / "rs_000003.dnd"), shapes, dtype)
    assert prec == np.dtype(dtype).itemsize
    assert t == pytest.approx(3 * w.dt, rel=1e-6)
    np.testing.assert_array_equal(np.concatenate([f.ravel() for f in fields]), first)
    cont = app(2, tmp_path / "b.bin", "restart=%s" % (tmp_path / "rs_000003.dnd"))
    full = app(5, tmp_path / "c.bin")
    np.testing.assert_array_equal(cont, full)


def test_dropin_iterator_feeder(gpu, tmp_path):
    """IteratorFeeder (iteratorfeeder.cpp:29-62): the iterator built with the
    code string of the jet grid's one patch group gets every patch and gives the
    same bits as the iterator the patches were added to directly; an iterator
    whose code string matches no group stops the run, naming the group's codes."""
    w = configs.jet(core=8, blocks=2)
    codes = "fx fy fz trans trans trans trans jet trans 0"
    direct = run_app(tmp_path, w, np.float32, 3, "hip")
    fed = run_app(tmp_path, w, np.float32, 3, "hip", extra=["codes=" + codes.replace(" ", ",")], tag="_fed")
    for a, b in zip(direct, fed):
        assert np.array_equal(a, b)
    grid, case = tmp_path / "grid", tmp_path / "case.bin"
    r = subprocess.run([APP, str(grid), str(case), str(tmp_path / "x.bin"), "1", repr(float(w.dt)), "hip",
                        "codes=fx,fy,fz,far,far,far,far,far,far,0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 5, r.stdout + r.stderr
    assert codes in r.stderr


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_viscous_operator(gpu, tmp_path, dtype):
    """HIP_CartesianIterator<5, EulerViscFlux<AusmPlus<...>, CompressibleViscFlux<5,
    PerfectGas>>> (the EaFlux composition) against the oracle with the viscous flux."""
    w = configs.viscous_shear(core=8, blocks=(2, 1, 1), periodic=False)
    og = oracle_after(w, dtype, 4)
    st = run_app(tmp_path, w, dtype, 4, "hip", extra=["visc"])
    err = rel_linf(st, og)
    print("drop-in viscous", np.dtype(dtype).name, err, flush=True)
    assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_vtk_output(gpu, tmp_path, dtype):
    """HIP_CartesianIterator::writeToVtk (PatchGrid::writeToVtk, patchgrid.cpp:673-695):
    one block per patch with the core cells; rho is the state's density as
    float32, the other quantities follow CompressibleVariables
    (compressiblevariables.h:81-105) of the same state, and the node points sit
    at origin + index * delta of the patch's core."""
    from vtkread import read_vtm, read_vts

    w = configs.jet(core=8, blocks=2)
    st = run_app(tmp_path, w, dtype, 3, "hip", extra=["vtk=%s" % (tmp_path / "flow")])
    blocks = read_vtm(str(tmp_path / "flow_000003.vtm"))
    assert [b[0] for b in blocks] == list(range(len(w.patches)))
    R, CV, CP, G = 287.2, 717.5, 1004.5, 1.4
    for p, P in enumerate(w.patches):
        assert blocks[p][1] == "flow_000003/flow_000003_%d.vts" % p
        r = read_vts(str(tmp_path / blocks[p][1]))
        s = P.seek
        core = np.asarray(st[p], dtype=np.float64).reshape(5, P.ni, P.nj, P.nk)[
            :, s[0]:P.ni - s[1], s[2]:P.nj - s[3], s[4]:P.nk - s[5]]
        core = np.transpose(core, (0, 3, 2, 1))            # VTK order: k, j, i
        rho = core[0]
        u, v, ww = core[1] / rho, core[2] / rho, core[3] / rho
        T = (core[4] / rho - 0.5 * (u * u + v * v + ww * ww)) / CV
        pr = rho * R * T
        np.testing.assert_array_equal(r["cells"]["rho"], rho.astype(np.float32))
        np.testing.assert_allclose(r["cells"]["p"], pr, rtol=2e-5)
        np.testing.assert_allclose(r["cells"]["T"], T, rtol=2e-5)
        np.testing.assert_allclose(r["cells"]["Ma"], np.sqrt((u * u + v * v + ww * ww) / (G * R * T)), rtol=2e-4,
                                   atol=1e-5)
        S = CP * np.log(T / 300.0) - R * np.log(pr / 1e5)
        np.testing.assert_allclose(r["cells"]["S"], S, rtol=2e-5)
        np.testing.assert_allclose(r["cells"]["U"], np.stack([u, v, ww], axis=-1), rtol=2e-4, atol=1e-5)
        h = [P.length[a] / n for a, n in enumerate((P.ni, P.nj, P.nk))]
        first = [P.origin[a] + s[2 * a] * h[a] for a in range(3)]
        np.testing.assert_allclose(r["points"][0, 0, 0], first, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(r["points"][-1, -1, -1],
                                   [P.origin[a] + (n - s[2 * a + 1]) * h[a]
                                    for a, n in enumerate((P.ni, P.nj, P.nk))], rtol=1e-6, atol=1e-7)


def cfl_numpy(w, states, dt, dtype):
    """CFL_max of the jet demo's diagnostics loop (applications/drnumJetDemo/
    main.h:196-205): max over cells of max(|u|, |u+a|, |u-a|) dt/dx, with the x
    velocity and spacing only, evaluated in FP64"""
    from drnum_amd import solver

    best = 0.0
    for P, U in zip(w.patches, states):
        U = U.astype(np.float64)
        u = U[1:4] / U[0]
        T = (U[4] / U[0] - 0.5 * (u * u).sum(axis=0)) / 717.5
        a = np.sqrt(1.4 * 287.2 * T)
        dx = float(solver.patch_spacing(P, dtype)[0])
        c = np.maximum(np.abs(u[0]), np.maximum(np.abs(u[0] + a), np.abs(u[0] - a))) * dt / dx
        best = max(best, float(c.max()))
    return best


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_cfl_time_step(gpu, tmp_path, dtype):
    """cfl=<target>: after every step dt *= target / CFL_max (the rescale of the
    jet demo, main.h:233).  The printed CFL of the last step is the CFL of the
    written state at the dt of that step, recomputed here with numpy; rk and hip
    modes take the same dt sequence and end in the same state."""
    w = configs.jet(core=8, blocks=2)
    outs = {}
    for m in ("rk", "hip"):
        log = []
        st = run_app(tmp_path, w, dtype, 4, m, extra=("cfl=0.3",), tag="cfl", stdout=log)
        lines = [ln.split() for ln in log[0].splitlines() if ln.startswith("dt ")]
        assert len(lines) == 4
        dts = [float(w.dt)] + [float(ln[1]) for ln in lines]
        cfls = [float(ln[3]) for ln in lines]
        for n in range(4):
            assert dts[n + 1] == pytest.approx(dts[n] * 0.3 / cfls[n], rel=1e-8)
        assert cfls[-1] == pytest.approx(cfl_numpy(w, st, dts[-2], dtype), rel=1e-5)
        outs[m] = (dts, st)
    assert outs["rk"][0] == outs["hip"][0]
    for a, b in zip(outs["rk"][1], outs["hip"][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dropin_cell_flags(gpu, tmp_path, dtype):
    """HostPatch::deactivate and HostPatch::setSplitFaces (patch_common.h:505,
    patch.h:895-931) through the drop-in: drnum_run deactivates the inactive
    cells of a flags file and gives the patches the faces between inside and
    outside cells, from which they fill their inside; every mode then equals
    the oracle with the same flags, and the filled inside equals the file's."""
    from test_gpu_cellflags import jet_flags

    w = configs.jet(core=8, blocks=2)
    flags = jet_flags(w)
    path = tmp_path / "flags.bin"
    path.write_bytes(b"".join(f.tobytes() for f in flags))
    og = O.OracleGrid(w, dtype, threads=8)
    og.init_state()
    og.compute_dependencies()
    for p, f in enumerate(flags):
        og.set_cell_flags(p, f)
    for _ in range(5):
        og.rk_step()
    out = []
    res = {m: run_app(tmp_path, w, dtype, 5, m, extra=("flags=%s" % path,), tag="fl", stdout=out)
           for m in ("rk", "hip", "graph")}
    want = "flags: %d inactive cells" % sum(int((f & 1).sum()) for f in flags)
    inside = "%d inside cells" % sum(int(((f >> 1) & 1).sum()) for f in flags)
    for text in out:
        assert want in text and inside in text, text
    for m, st in res.items():
        err = rel_linf(st, og)
        print("drop-in flags", m, np.dtype(dtype).name, err, flush=True)
        assert err <= TOL[dtype]
    for m in ("hip", "graph"):
        for a, b in zip(res["rk"], res[m]):
            assert np.array_equal(a, b), m
