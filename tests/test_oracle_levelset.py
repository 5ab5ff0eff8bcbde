"""Pins the level-set classification, mirror tables and slip fill (CPU only).

Oracle (levelsetobject.cpp:44-388, cartesianpatch.cpp:1512-1657,
cpu_levelsetobjectbc.h:58-90, lsobccompressibleeulerop.h:49-144) against:

  * G = min_s(|x - c_s| - r_s) evaluated independently;
  * layer 0 (no min-distance replacement) = cells with G < 0 that have a face
    neighbour across the zero level (brute force over all cells);
  * layer L = inner cells at face-graph distance L from layer 0;
  * min-distance replacement: every kept cell is deep enough, every
    replacement is a shallower cell's face neighbour with smaller G;
  * normals are unit vectors along the exact sphere normal (to O(h^2));
  * the mirror point x - 2 G n lies across the surface (G(mirror) ~ -G) and
    the mirror weights sum to one and reproduce it;
  * the slip fill: rest state preserved; flow normal to a wall is stopped
    (relax 0.5); flow tangential to it passes unchanged.

And the product's host-side builder (libdrnum.so) against the oracle, bit
for bit, in FP64 and FP32 geometry.
"""
import numpy as np
import pytest

from drnum_amd import configs
from oracle import oracle as O

from refsol import R, ijk


def small_ls(blocks=(2, 1, 1), core=16, n_spheres=2, seed=7, r_range=(2.5, 3.5)):
    return configs.levelset_spheres(core=core, blocks=blocks, n_spheres=n_spheres, seed=seed, r_range=r_range)


def sphere_g(w, p):
    x, y, z = w.cell_centres(p)
    X = np.stack([x, y, z])
    return np.min([np.linalg.norm(X - s[:3, None], axis=0) - s[3] for s in w.spheres], axis=0)


def face_neighbours(P, c):
    i, j, k = (int(v[0]) for v in ijk(P, [c]))
    out = []
    for (a, b, d) in ((i, j, k - 1), (i, j, k + 1), (i, j - 1, k), (i, j + 1, k), (i - 1, j, k), (i + 1, j, k)):
        if 0 <= a < P.ni and 0 <= b < P.nj and 0 <= d < P.nk:
            out.append((a * P.nj + b) * P.nk + d)
    return out


def oracle_ls(w, dtype=np.float64, n_inner=3, min_rel=0.1):
    g = O.OracleGrid(w, dtype)
    g.levelset(w.spheres, n_inner=n_inner, n_outer=0, min_inner_rel=min_rel)
    return g


def test_levelset_values():
    w = small_ls()
    g = oracle_ls(w)
    for p in range(len(w.patches)):
        np.testing.assert_allclose(g.levelset_g(p), sphere_g(w, p), rtol=0, atol=1e-14)


@pytest.mark.parametrize("blocks", [(1, 1, 1), (2, 1, 1)])
def test_layer0_without_replacement_is_the_sign_change_band(blocks):
    """On one patch every mirror point has a donor octet, so the list is the
    whole band; with seek layers, cells whose mirror point leaves the donor
    zone are dropped (levelsetobject.cpp:344-359) and the rest keep their order."""
    w = small_ls(blocks=blocks)
    g = oracle_ls(w, n_inner=1, min_rel=0.0)
    for p, P in enumerate(w.patches):
        G = g.levelset_g(p)
        band = (G.max() - G.min()) * 1e-7
        want = []
        for c in range(P.cells):
            if G[c] < 0 and any((G[n] + band) * (G[c] - band) < 0 or (G[n] - band) * (G[c] + band) < 0
                                for n in face_neighbours(P, c)):
                want.append(c)
        got = g.ls_cells(p, 0)["cells"].tolist()
        if blocks == (1, 1, 1):
            assert got == want and len(want) > 0
        else:
            pos = [want.index(c) for c in got]
            assert pos == sorted(pos)


def graph_distance(P, G, seeds, max_d):
    """Face-graph distance through inner cells (G < 0) from a seed set."""
    dist = {int(c): 0 for c in seeds}
    front = [int(c) for c in seeds]
    for d in range(1, max_d + 1):
        nxt = []
        for c in front:
            for n in face_neighbours(P, c):
                if n not in dist and G[n] < 0:
                    dist[n] = d
                    nxt.append(n)
        front = nxt
    return dist


@pytest.mark.parametrize("min_rel", [0.0, 0.1])
def test_deeper_layers_are_graph_distance_shells(min_rel):
    """One patch, so no list entry is dropped by the mirror search: layer L is
    exactly the set of inner cells at face-graph distance L from layer 0."""
    w = small_ls(blocks=(1, 1, 1))
    g = oracle_ls(w, n_inner=3, min_rel=min_rel)
    P = w.patches[0]
    G = g.levelset_g(0)
    layers = [g.ls_cells(0, L)["cells"] for L in range(3)]
    dist = graph_distance(P, G, layers[0], 2)
    for L in (1, 2):
        want = sorted(c for c, d in dist.items() if d == L)
        assert sorted(layers[L].tolist()) == want and len(want) > 0
    allc = np.concatenate(layers)
    assert len(np.unique(allc)) == len(allc)


def test_min_distance_replacement():
    """levelsetobject.cpp:164-222: a layer-0 cell whose G is above
    -min_rel * (max - min of its face neighbours' G)/2 is replaced by those
    neighbours with smaller G."""
    w = small_ls(blocks=(1, 1, 1))
    raw = oracle_ls(w, n_inner=1, min_rel=0.0)
    rep = oracle_ls(w, n_inner=1, min_rel=0.1)
    for p, P in enumerate(w.patches):
        G = raw.levelset_g(p)
        before = set(raw.ls_cells(p, 0)["cells"].tolist())
        after = rep.ls_cells(p, 0)["cells"].tolist()
        assert len(after) == len(set(after))

        def shallow(c):
            nb = face_neighbours(P, c)
            gd = 0.5 * (max(G[n] for n in nb) - min(G[n] for n in nb))
            return G[c] > -0.1 * gd

        for c in after:
            if c in before:
                assert not shallow(c)
            else:
                # a replacement: a face neighbour with smaller G of some removed shallow cell
                assert any(c in face_neighbours(P, r) and G[c] < G[r] and shallow(r) for r in before)
        removed = [c for c in before if c not in after]
        assert all(shallow(c) for c in removed)


def test_normals_and_mirror_points():
    w = small_ls(n_spheres=1, r_range=(3.0, 3.0))
    g = oracle_ls(w)
    s = w.spheres[0]
    for p, P in enumerate(w.patches):
        X = np.stack(w.cell_centres(p))
        for L in range(3):
            d = g.ls_cells(p, L)
            if len(d["cells"]) == 0:
                continue
            n = d["g"][:, 1:4]
            np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, rtol=0, atol=1e-14)
            x = X[:, d["cells"]]
            exact = (x - s[:3, None]) / np.linalg.norm(x - s[:3, None], axis=0)
            cosang = np.sum(n.T * exact, axis=0)
            assert cosang.min() > 0.98
            # mirror point: reflection of the cell centre across the zero level
            m = x - 2.0 * d["g"][:, 0] * n.T
            gm = np.linalg.norm(m - s[:3, None], axis=0) - s[3]
            h = P.dx(0)
            np.testing.assert_allclose(gm, -d["g"][:, 0], rtol=0, atol=0.1 * h)
            # the weights sum to one and reproduce the mirror point (local frame)
            np.testing.assert_allclose(d["weights"].sum(axis=1), 1.0, rtol=0, atol=8e-16)
            xl = X - np.array(P.origin)[:, None]
            rep = np.einsum("cs,acs->ac", d["weights"], xl[:, d["donors"]])
            np.testing.assert_allclose(rep, m - np.array(P.origin)[:, None], rtol=0, atol=1e-3 * h)


def plane_wall_workload(u, v):
    """A sphere so large that it acts as the plane x = 0.3 (solid for x < 0.3)."""
    h = 1.0 / 16
    P = configs.Patch(16, 16, 16, [0] * 6, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [configs.BC_SLIP] * 6,
                      np.zeros((6, configs.NPAR)), [0, 0, 0])
    big = 1.0e4
    spheres = np.array([[0.3 - big, 0.5, 0.5, -0.0]])
    # G = |x - c| - r < 0 inside; with c far to the -x side and r = big, G ~ (x - 0.3) near the plane
    spheres[0, 3] = big

    def init(x, y, z):
        one = np.ones_like(x)
        return np.stack([one * 1.0, one / R, u * one, v * one, 0 * x])

    w = configs.Workload("plane", [P], configs.RECON_UPWIND2, configs.LIM_VANALBADA, [0.0] * 3, [0.25, 0.5, 1.0],
                         1e-3, init, spheres=spheres)
    return w, h


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fill_rest_state_is_preserved(dtype):
    w, _ = plane_wall_workload(0.0, 0.0)
    g = oracle_ls(w, dtype)
    g.init_state()
    U0 = g.field(0, 0).copy()
    g.levelset_fill()
    np.testing.assert_allclose(g.field(0, 0), U0, rtol=8 * np.finfo(dtype).eps, atol=0)


def test_fill_stops_normal_flow_and_keeps_tangential_flow():
    cells = None
    for u, v in ((0.4, 0.0), (0.0, 0.4)):
        w, h = plane_wall_workload(u, v)
        g = oracle_ls(w, np.float64)
        g.init_state()
        U0 = g.field(0, 0).copy()
        g.levelset_fill()
        U = g.field(0, 0)
        cells = np.concatenate([g.ls_cells(0, L)["cells"] for L in range(3)])
        assert len(cells) > 0
        np.testing.assert_allclose(U[0, cells], U0[0, cells], rtol=1e-14)
        np.testing.assert_allclose(U[4, cells], U0[4, cells], rtol=1e-14)
        if u:
            # normal incidence: reflected (-u) blended half-and-half with the current (+u) value
            assert np.abs(U[1, cells]).max() < 1e-3 * u
        else:
            np.testing.assert_allclose(U[2, cells], U0[2, cells], rtol=1e-3)
            assert np.abs(U[1, cells]).max() < 1e-3 * v
        # cells outside the lists are untouched
        mask = np.ones(U.shape[1], bool)
        mask[cells] = False
        np.testing.assert_array_equal(U[:, mask], U0[:, mask])


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("case", ["two", "four_blocks"])
def test_host_levelset_tables_match_oracle_bitwise(fp64, case):
    from drnum_amd import host

    w = small_ls() if case == "two" else small_ls(blocks=(2, 2, 1), n_spheres=3)
    dtype = np.float64 if fp64 else np.float32
    g = oracle_ls(w, dtype)
    hg = host.HostGrid.from_workload(w, fp64_geometry=fp64)
    hg.levelset_spheres(w.spheres, n_inner=3, n_outer=0, min_inner_rel=0.1)
    for p, P in enumerate(w.patches):
        np.testing.assert_array_equal(hg.levelset_values(p, P.cells).astype(dtype), g.levelset_g(p))
        for L in range(3):
            a = g.ls_cells(p, L)
            b = hg.ls_cells(p, L)
            np.testing.assert_array_equal(b["cells"], a["cells"])
            np.testing.assert_array_equal(b["donors"], a["donors"])
            np.testing.assert_array_equal(b["g"].astype(dtype), a["g"])
            np.testing.assert_array_equal(b["weights"].astype(dtype), a["weights"])


# ---- the simple-wall operator (LSOBCCompressibleSWallOp) --------------------

def linear_wall_workload(u_of, v_of):
    """The plane wall of plane_wall_workload with velocities u_of(x, y), v_of(x, y)."""
    w, h = plane_wall_workload(0.0, 0.0)

    def init(x, y, z):
        one = np.ones_like(x)
        return np.stack([one * 1.0, one / R, u_of(x, y), v_of(x, y), 0 * x])

    w.init = init
    return w, h


def test_swall_fill_stops_normal_and_tangential_flow():
    """The simple wall reverses the whole mirrored momentum and blends it half
    and half with the current one (lsobccompressibleswallop.h:103-106): in a
    uniform flow every layer cell comes to rest, where the slip wall keeps
    the tangential flow.  Density and total energy come from the mirror sum."""
    for u, v in ((0.4, 0.0), (0.0, 0.4), (0.3, -0.2)):
        w, h = plane_wall_workload(u, v)
        g = oracle_ls(w, np.float64)
        g.set_levelset_op(1)
        g.init_state()
        U0 = g.field(0, 0).copy()
        g.levelset_fill()
        U = g.field(0, 0)
        cells = np.concatenate([g.ls_cells(0, L)["cells"] for L in range(3)])
        assert len(cells) > 0
        np.testing.assert_allclose(U[0, cells], U0[0, cells], rtol=1e-14)
        np.testing.assert_allclose(U[4, cells], U0[4, cells], rtol=1e-14)
        assert np.abs(U[1:4, cells]).max() < 1e-12
        mask = np.ones(U.shape[1], bool)
        mask[cells] = False
        np.testing.assert_array_equal(U[:, mask], U0[:, mask])


def test_swall_fill_of_linear_momentum_is_the_mirror_formula():
    """Linear momentum fields are reproduced by the trilinear mirror weights,
    so each layer cell must hold 1/2 (-m(x_m)) + 1/2 m(x_c), with the mirror
    point x_m = x_c - 2 G n of the cell (levelsetobject.cpp:300-330)."""
    u_of = lambda x, y: 0.2 + 0.3 * y - 0.1 * x
    v_of = lambda x, y: -0.1 + 0.2 * x
    w, h = linear_wall_workload(u_of, v_of)
    g = oracle_ls(w, np.float64)
    g.set_levelset_op(1)
    g.init_state()
    g.levelset_fill()
    U = g.field(0, 0)
    P = w.patches[0]
    checked = 0
    for L in range(3):
        t = g.ls_cells(0, L)
        i, j, k = ijk(P, t["cells"])
        xc = np.array(P.origin)[:, None] + (np.stack([i, j, k]) + 0.5) * h
        G, n = t["g"][:, 0], t["g"][:, 1:4].T
        xm = xc - 2.0 * G * n
        full = np.abs(t["weights"].sum(axis=1) - 1.0) < 1e-12
        for q, f in ((1, u_of), (2, v_of)):
            expect = 0.5 * (-f(xm[0], xm[1])) + 0.5 * f(xc[0], xc[1])
            np.testing.assert_allclose(U[q, t["cells"]][full], expect[full], rtol=0, atol=2e-3 * 0.3 * h)
        checked += int(full.sum())
    assert checked > 50


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_swall_and_slip_differ_only_in_momentum(dtype):
    w, _ = plane_wall_workload(0.3, -0.2)
    out = {}
    for op in (0, 1):
        g = oracle_ls(w, dtype)
        g.set_levelset_op(op)
        g.init_state()
        g.levelset_fill()
        out[op] = g.field(0, 0).copy()
    np.testing.assert_array_equal(out[0][[0, 4]], out[1][[0, 4]])
    assert np.abs(out[0][1:4] - out[1][1:4]).max() > 0.05
