"""CPU checks of the C ABI and the host-side logic (no GPU calls).

  * both libraries load and export every function that include/*.h declares,
    and the Python bindings bind exactly that set;
  * the host grid reader parses the reference's grid-file format
    (patchgrid.cpp:479-561, the writer in src/Python/drnum.py:199-244) into
    the same patches, and the same donor tables, as the programmatic set-up;
  * error paths return codes and messages instead of stopping the process;
  * the benchmark workload has the stated size.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from drnum_amd import configs, hip, host

from casefiles import grid_file_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(drnh?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib_path,bound", [
    ("drnhip.h", hip.LIB_PATH, lambda: hip._SIGS),
    ("drnum_host.h", host.LIB_PATH, lambda: host._SIGS),
])
def test_library_exports_every_declared_symbol(header, lib_path, bound):
    names = declared(header)
    assert len(names) > 20
    L = C.CDLL(lib_path)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(bound()) == names


def test_hip_library_is_a_gfx950_code_object_host():
    """The HIP library carries device code for gfx950 (offload bundle section)."""
    data = open(hip.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_grid_reader_matches_programmatic_setup(tmp_path):
    w = configs.jet(core=6, blocks=2)
    path = tmp_path / "grid"
    path.write_text(grid_file_text(w))
    a = host.HostGrid(True)
    a.read(str(path))
    b = host.HostGrid.from_workload(w)
    assert a.num_patches() == len(w.patches)
    for p, P in enumerate(w.patches):
        ia, ib = a.patch_info(p), b.patch_info(p)
        assert ia == ib
        assert ia["n"] == [P.ni, P.nj, P.nk] and ia["seek"] == list(P.seek)
    a.compute_dependencies()
    b.compute_dependencies()
    for p in range(len(w.patches)):
        ta, tb = a.tables(p), b.tables(p)
        for k in ("receivers", "rec_concat", "donor_index", "donor_weight"):
            np.testing.assert_array_equal(ta[k], tb[k])
        assert ta["donors"] == tb["donors"]


def test_grid_reader_scale(tmp_path):
    w = configs.jet(core=6, blocks=2)
    path = tmp_path / "grid"
    path.write_text(grid_file_text(w))
    a = host.HostGrid(True)
    a.read(str(path), 2.0)
    info = a.patch_info(1)
    P = w.patches[1]
    np.testing.assert_allclose(info["origin"], 2.0 * np.array(P.origin), rtol=1e-15)
    np.testing.assert_allclose(info["length"], 2.0 * np.array(P.length), rtol=1e-15)


def test_grid_reader_errors_are_reported(tmp_path):
    g = host.HostGrid(True)
    with pytest.raises(RuntimeError, match="cannot open"):
        g.read(str(tmp_path / "nope"))
    # a turned patch is read with its base vectors; parallel or zero ones are refused
    rot = tmp_path / "rot"
    rot.write_text("1001 { 0 0 0  0 1 0  -1 0 0  1  4 4 4  0 0 0 0 0 0  1 1 1 }\n0\n")
    g.read(str(rot))
    bi, bj = g.patch_base(0)
    assert bi.tolist() == [0.0, 1.0, 0.0] and bj.tolist() == [-1.0, 0.0, 0.0]
    A, b = g.patch_transform(0)
    assert np.array_equal(A, [[0, -1, 0], [1, 0, 0], [0, 0, 1]]) and np.array_equal(b, [0, 0, 0])
    for base in ("2 0 0  4 0 0", "0 0 0  0 1 0"):
        bad = tmp_path / "degenerate"
        bad.write_text("1001 { 0 0 0  %s  1  4 4 4  0 0 0 0 0 0  1 1 1 }\n0\n" % base)
        with pytest.raises(RuntimeError, match="zero or parallel"):
            host.HostGrid(True).read(str(bad))
    bad2 = tmp_path / "type"
    bad2.write_text("1002 { }\n0\n")
    with pytest.raises(RuntimeError, match="unsupported patch type"):
        g.read(str(bad2))


def test_host_rejects_bad_patch_ids():
    g = host.HostGrid.from_workload(configs.jet(core=6, blocks=2))
    info = np.zeros(9, np.int32)
    od = np.zeros(9, np.float64)
    assert g.L.drnh_grid_patch_info(g.h, 99, info.ctypes.data_as(C.POINTER(C.c_int32)),
                                    od.ctypes.data_as(C.POINTER(C.c_double))) < 0
    assert g.L.drnh_grid_patch_info(g.h, -1, info.ctypes.data_as(C.POINTER(C.c_int32)),
                                    od.ctypes.data_as(C.POINTER(C.c_double))) < 0


def test_min_ch_length():
    w = configs.jet(core=6, blocks=2)
    g = host.HostGrid.from_workload(w)
    assert g.min_ch_length() == pytest.approx(1.0 / 12, rel=1e-15)


def test_jet_benchmark_workload_sizes():
    w = configs.jet(core=64, blocks=4)
    assert len(w.patches) == 64
    assert w.total_cells == 19_248_832
    assert w.core_cells == 16_777_216
    # receivers per exchange: every stored cell outside the cores
    assert w.total_cells - w.core_cells == 2_471_616
    dims = sorted({(P.ni, P.nj, P.nk) for P in w.patches})
    assert {d for dd in dims for d in dd} == {66, 68}


def test_periodic_box_every_patch_has_seek_layers():
    w = configs.kelvin_helmholtz(core=8, blocks=2)
    for P in w.patches:
        assert P.seek == [2] * 6 and (P.ni, P.nj, P.nk) == (12, 12, 12)


def test_bench_reads_the_committed_traffic_figure():
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    # the summary holds the tiling the autotuner picked when it was taken
    figs = [t for t in (b.measured_traffic(c) for c in range(64)) if t is not None]
    assert figs and all(0.9e9 < t < 1.2e9 for t in figs)
    assert b.profile_hash(b.TRAFFIC_TABLE) is not None
    assert b.measured_traffic(999) is None


def test_jet_exchange_bytes_follow_from_the_grid():
    """Every receiver of the benchmark jet is a seek cell with exactly one
    unit-weight donor, so one FP32 exchange moves 52 B per receiver (index and
    5 values written, index, weight and 5 values read): the bench's figure.
    With the z-face exchange fused into the stage kernel, the exchange keeps
    the receivers outside the z-face seek layers of the x/y core columns."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    w = configs.jet(core=64, blocks=4)
    hg = host.HostGrid.from_workload(w, fp64_geometry=True)
    hg.compute_dependencies(w.periodic)
    n_rec = n_pairs = n_zface = 0
    for p, P in enumerate(w.patches):
        t = hg.tables(p)
        n_rec += len(t["receivers"])
        n_pairs += len(t["donor_index"])
        assert np.all(t["donor_weight"] == 1.0)
        r = t["receivers"]
        i, j, k = r // (P.nj * P.nk), (r // P.nk) % P.nj, r % P.nk
        s = P.seek
        core_xy = (i >= s[0]) & (i < P.ni - s[1]) & (j >= s[2]) & (j < P.nj - s[3])
        n_zface += int(np.sum(core_xy & ((k < s[4]) | (k >= P.nk - s[5]))))
    assert n_rec == n_pairs == w.total_cells - w.core_cells
    ex, ls = b.exchange_fill_bytes(hg, w, 4)
    assert ex == 52 * n_rec == 128_524_032 and ls == 0
    # 48 interior z faces, both sides, 64 x 64 core columns, 2 layers
    assert n_zface == 48 * 2 * 64 * 64 * 2 == 786_432
    assert 52 * (n_rec - n_zface) == 87_629_568   # the exchange after a fused stage
    assert b.measured_exchange_traffic() == pytest.approx(110.7e6)


def test_dropin_app_reports_missing_device(tmp_path):
    """Without a GPU the C++ application stops with the device error of
    drn_ctx_create (the C ABI's return code, turned into drnum::DeviceError)."""
    import subprocess

    from casefiles import case_file_bytes

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    w = configs.jet(core=6, blocks=2)
    (tmp_path / "grid").write_text(grid_file_text(w))
    (tmp_path / "case").write_bytes(case_file_bytes(w, np.float32))
    app = os.path.join(ROOT, "drnum_amd", "lib", "drnum_run")
    r = subprocess.run([app, str(tmp_path / "grid"), str(tmp_path / "case"), str(tmp_path / "out"), "1", "1e-4", "rk"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 4, r.stdout + r.stderr
    assert "drn_ctx_create" in r.stderr
    assert "grid: 8 patches read" in r.stdout


# ---------------------------------------------------------------------------
# restart files (.dnd, patchgrid.cpp:563-610)
# ---------------------------------------------------------------------------

def _fields(dtype, seed=4):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((5, n)).astype(dtype) for n in (7, 12, 3)]


@pytest.mark.parametrize("dtype,code,fmt", [(np.float32, 4, ">f4"), (np.float64, 8, ">f8")])
def test_dnd_bytes_are_the_reference_stream_layout(tmp_path, dtype, code, fmt):
    """QDataStream is big-endian: int32 precision, the time as `real`, then each
    patch's field [var][cell] (patchgrid.cpp:563-583, patch.cpp:776-782)."""
    fields = _fields(dtype)
    path = str(tmp_path / "run_000010.dnd")
    host.dnd_write(path, fields, 0.125)
    want = np.array([code], dtype=">i4").tobytes() + np.array([0.125], dtype=fmt).tobytes()
    want += b"".join(f.astype(fmt).tobytes() for f in fields)
    assert open(path, "rb").read() == want


@pytest.mark.parametrize("wdt,rdt", [(np.float32, np.float32), (np.float64, np.float64), (np.float32, np.float64),
                                     (np.float64, np.float32)])
def test_dnd_round_trip_and_conversion(tmp_path, wdt, rdt):
    fields = _fields(wdt)
    base = str(tmp_path / "state")
    host.dnd_write(base + ".dnd", fields, 2.5)
    got, t, prec = host.dnd_read(base, [f.shape for f in fields], rdt)   # ".dnd" appended when missing
    assert t == 2.5 and prec == (4 if wdt == np.float32 else 8)
    for a, b in zip(got, fields):
        np.testing.assert_array_equal(a, b.astype(rdt))


def test_dnd_rejects_a_file_of_another_grid(tmp_path):
    fields = _fields(np.float32)
    path = str(tmp_path / "x.dnd")
    host.dnd_write(path, fields, 1.0)
    with pytest.raises(RuntimeError):
        host.dnd_read(path, [f.shape for f in fields[:2]], np.float32)         # file longer than the grid
    with pytest.raises(RuntimeError):
        host.dnd_read(path, [f.shape for f in fields] + [(5, 4)], np.float32)  # file shorter than the grid
    with pytest.raises(RuntimeError):
        host.dnd_read(str(tmp_path / "missing.dnd"), [(5, 1)], np.float32)


def test_dnd_rejects_a_corrupted_file(tmp_path):
    """A precision word other than 4 or 8, a file cut inside its header and a
    file cut inside its last value are refused, not read as a state."""
    fields = _fields(np.float32)
    shapes = [f.shape for f in fields]
    path = tmp_path / "good.dnd"
    host.dnd_write(str(path), fields, 1.0)
    raw = path.read_bytes()
    bad = tmp_path / "bad.dnd"
    for data in (np.array([7], dtype=">i4").tobytes() + raw[4:], raw[:6], raw[:len(raw) - 3]):
        bad.write_bytes(data)
        with pytest.raises(RuntimeError):
            host.dnd_read(str(bad), shapes, np.float32)
    got, t, prec = host.dnd_read(str(path), shapes, np.float32)
    assert t == 1.0 and prec == 4
