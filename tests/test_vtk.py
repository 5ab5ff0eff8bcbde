"""VTK output (PatchGrid::writeToVtk, patchgrid.cpp:673-695; CartesianPatch::
createVtkDataSet, cartesianpatch.cpp:1661-1730): the multiblock index and the
per-patch structured grids of core cells, read back without the VTK library."""
import os

import numpy as np

from drnum_amd import host
from vtkread import read_vtm, read_vts


def test_multiblock_index(tmp_path):
    base = str(tmp_path / "flow")
    d = host.vtk_write_multiblock(base, 55, 3)
    assert d == base + "_000055"
    assert os.path.isdir(d)
    assert read_vtm(d + ".vtm") == [(p, "flow_000055/flow_000055_%d.vts" % p) for p in range(3)]
    d2 = host.vtk_write_multiblock(base, -1, 1)
    assert d2 == base and read_vtm(base + ".vtm") == [(0, "flow/flow_0.vts")]


def test_patch_block_holds_the_core_cells_in_vtk_order(tmp_path):
    ni, nj, nk = 7, 6, 5
    seek = [2, 0, 1, 2, 0, 1]            # seek layers per face xm, xp, ym, yp, zm, zp
    origin, delta = [0.5, -1.0, 2.0], [0.25, 0.5, 0.125]
    N = ni * nj * nk
    cell = np.arange(N, dtype=np.float64)
    i, j, k = np.unravel_index(np.arange(N), (ni, nj, nk))   # cell = i*nj*nk + j*nk + k
    scalars = {"a": 1000.0 * i + 100.0 * j + k, "b": -cell}
    vectors = {"U": np.stack([i + 0.5, j + 0.25, k + 0.125])}
    path = str(tmp_path / "p.vts")
    host.vtk_write_patch(path, (ni, nj, nk), seek, origin, delta, scalars, vectors)
    r = read_vts(path)
    ci, cj, ck = ni - 2, nj - 3, nk - 1
    assert r["extent"] == [0, ci, 0, cj, 0, ck]
    assert list(r["cells"]) == ["a", "b", "U"]
    # cell (kk, jj, ii) of the block is patch cell (i0 + ii, j0 + jj, k0 + kk)
    for kk in range(ck):
        for jj in range(cj):
            for ii in range(ci):
                I, J, K = 2 + ii, 1 + jj, 0 + kk
                c = I * nj * nk + J * nk + K
                assert r["cells"]["a"][kk, jj, ii] == 1000.0 * I + 100.0 * J + K
                assert r["cells"]["b"][kk, jj, ii] == -c
                np.testing.assert_array_equal(r["cells"]["U"][kk, jj, ii], [I + 0.5, J + 0.25, K + 0.125])
    # node points of the core: origin + index * delta, node (0,0,0) at the first core cell's corner
    P = r["points"]
    assert P.shape == (ck + 1, cj + 1, ci + 1, 3)
    np.testing.assert_allclose(P[0, 0, 0], [0.5 + 2 * 0.25, -1.0 + 1 * 0.5, 2.0], rtol=0, atol=1e-6)
    np.testing.assert_allclose(P[-1, -1, -1], [0.5 + 7 * 0.25, -1.0 + 4 * 0.5, 2.0 + 4 * 0.125], rtol=0, atol=1e-6)
    np.testing.assert_allclose(P[1, 2, 3] - P[0, 0, 0], [3 * 0.25, 2 * 0.5, 1 * 0.125], rtol=0, atol=1e-6)
