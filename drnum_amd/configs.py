"""Synthetic patch-grid workloads (the benchmark and test cases of this path).

A workload is a plain description: a list of patch dicts in the reference's
grid-file vocabulary (patchgrid.cpp:479-561, cartesianpatch.cpp:98-133:
origin, NI NJ NK, six seek-layer counts, lengths) plus the face functor of
each patch face (the xWallM/xWallP/... surface), the time step, the RK
coefficients and the initial state.  Nothing here computes on a device.

Face functor codes match ``include/drnhip.h`` (DRN_BC_*).
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import gas

BC_SLIP, BC_FARFIELD, BC_TRANSMISSIVE, BC_JET, BC_NONE = 0, 1, 2, 3, 4
RECON_UPWIND1, RECON_UPWIND2 = 0, 1
LIM_VANALBADA, LIM_MINMOD, LIM_SECONDORDER = 0, 1, 2
NPAR = 16


@dataclasses.dataclass
class Patch:
    ni: int
    nj: int
    nk: int
    seek: List[int]            # imin, imax, jmin, jmax, kmin, kmax
    origin: List[float]        # lower corner of local cell (0,0,0)
    length: List[float]        # Lx, Ly, Lz
    bc: List[int]              # face functor: xm, xp, ym, yp, zm, zp
    bc_par: np.ndarray         # (6, NPAR) float64
    gidx: List[int]            # global index of local cell (0,0,0) (for lattice grids)
    # base vectors of the local i and j axes in the inertial frame, as a grid
    # file gives them (patch.cpp:63-85); the default is axis aligned
    base_i: List[float] = dataclasses.field(default_factory=lambda: [1.0, 0.0, 0.0])
    base_j: List[float] = dataclasses.field(default_factory=lambda: [0.0, 1.0, 0.0])

    @property
    def aligned(self) -> bool:
        return list(self.base_i) == [1.0, 0.0, 0.0] and list(self.base_j) == [0.0, 1.0, 0.0]

    def rotation(self) -> np.ndarray:
        """Local -> inertial rotation (columns: unit i, j, k axes), the matrix of
        CoordTransform::setMatrixFromBaseIJ (coordtransform.cpp:37-58) in float64."""
        bi = np.asarray(self.base_i, dtype=np.float64)
        bj = np.asarray(self.base_j, dtype=np.float64)
        bi = bi / np.linalg.norm(bi)
        bj = bj / np.linalg.norm(bj)
        bk = np.cross(bi, bj)
        bj = bj - (bi @ bj) * bi
        bj = bj / np.linalg.norm(bj)
        bk = bk / np.linalg.norm(bk)
        return np.stack([bi, bj, bk], axis=1)

    @property
    def cells(self) -> int:
        return self.ni * self.nj * self.nk

    @property
    def core_cells(self) -> int:
        s = self.seek
        return (self.ni - s[0] - s[1]) * (self.nj - s[2] - s[3]) * (self.nk - s[4] - s[5])

    def dx(self, axis: int) -> float:
        return self.length[axis] / [self.ni, self.nj, self.nk][axis]


@dataclasses.dataclass
class Workload:
    name: str
    patches: List[Patch]
    recon: int
    limiter: int
    periodic: List[float]      # period per axis, 0 = none
    alpha: List[float]
    dt: float
    init: Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray]  # (x,y,z) -> (5, n) primitive p,T,u,v,w
    spheres: Optional[np.ndarray] = None   # (n, 4) xc, yc, zc, r for the level-set case
    ls_inner_layers: int = 3
    ls_min_inner_rel: float = 0.1
    steps: int = 1
    viscous: bool = False   # CompressibleViscFlux added to every interior face flux

    @property
    def total_cells(self) -> int:
        return sum(p.cells for p in self.patches)

    @property
    def core_cells(self) -> int:
        return sum(p.core_cells for p in self.patches)

    def cell_centres(self, p: int):
        """Inertial cell-centre coordinates (float64) of patch p, flattened i,j,k
        with k fastest: origin + R (local centre) for a rotated patch."""
        P = self.patches[p]
        if P.aligned:
            xs = [P.origin[a] + (np.arange(n) + 0.5) * P.dx(a) for a, n in enumerate((P.ni, P.nj, P.nk))]
            X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
            return X.ravel(), Y.ravel(), Z.ravel()
        xs = [(np.arange(n) + 0.5) * P.dx(a) for a, n in enumerate((P.ni, P.nj, P.nk))]
        X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
        loc = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
        g = P.rotation() @ loc + np.asarray(P.origin, dtype=np.float64)[:, None]
        return g[0], g[1], g[2]

    def initial_state(self, p: int, dtype) -> np.ndarray:
        """Conserved variables (5, N) of patch p in `dtype`, converted with the
        reference's primitiveToConservative (perfectgas.h:56-64).  A rotated
        patch holds its velocity in its own frame: R^T u."""
        P = self.patches[p]
        x, y, z = self.cell_centres(p)
        prim = self.init(x, y, z)
        if not P.aligned:
            prim = np.array(prim, dtype=np.float64, copy=True)
            prim[2:5] = P.rotation().T @ prim[2:5]
        return gas.prim_to_cons(prim[0], prim[1], prim[2], prim[3], prim[4], dtype)


def _par() -> np.ndarray:
    return np.zeros((6, NPAR), dtype=np.float64)


def block_grid(ncells: Sequence[int], blocks: Sequence[int], length: Sequence[float], seek: int,
               periodic: Sequence[bool], domain_bc: Sequence[int], interior_bc: int = BC_TRANSMISSIVE,
               domain_par: Optional[np.ndarray] = None) -> List[Patch]:
    """A box of ncells cells cut into blocks[0]*blocks[1]*blocks[2] patches of equal
    core size.  Every face with a neighbour (or a periodic image) carries `seek`
    seek layers; domain faces carry none and get `domain_bc`."""
    core = [ncells[a] // blocks[a] for a in range(3)]
    assert all(core[a] * blocks[a] == ncells[a] for a in range(3))
    h = [length[a] / ncells[a] for a in range(3)]
    patches = []
    for bi in range(blocks[0]):
        for bj in range(blocks[1]):
            for bk in range(blocks[2]):
                b = (bi, bj, bk)
                sk, n, org, L, bc, gidx = [], [], [], [], [], []
                par = _par()
                for a in range(3):
                    lo = seek if (b[a] > 0 or periodic[a]) else 0
                    hi = seek if (b[a] < blocks[a] - 1 or periodic[a]) else 0
                    sk += [lo, hi]
                    n.append(core[a] + lo + hi)
                    g0 = b[a] * core[a] - lo
                    gidx.append(g0)
                    org.append(g0 * h[a])
                    L.append(n[-1] * h[a])
                    for side, s in ((0, lo), (1, hi)):
                        f = 2 * a + side
                        if s > 0:
                            bc.append(interior_bc)
                        else:
                            bc.append(domain_bc[f])
                            if domain_par is not None:
                                par[f] = domain_par[f]
                patches.append(Patch(n[0], n[1], n[2], sk, org, L, bc, par, gidx))
    return patches


# ----------------------------------------------------------------------------
# Config 1: Sod shock tube, one 64^3 patch
# ----------------------------------------------------------------------------

def sod(n: int = 64, steps: int = 100) -> Workload:
    """Sod along x on one n^3 patch, FP64 reference config.  y/z faces use the
    slip functor: for this y,z-uniform state with v=w=0 it gives the same face
    fluxes as periodicity.  x faces are transmissive."""
    bc = [BC_TRANSMISSIVE, BC_TRANSMISSIVE, BC_SLIP, BC_SLIP, BC_SLIP, BC_SLIP]
    patches = [Patch(n, n, n, [0] * 6, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], bc, _par(), [0, 0, 0])]

    def init(x, y, z):
        left = x < 0.5
        rho = np.where(left, 1.0, 0.125)
        p = np.where(left, 1.0, 0.1)
        T = p / (rho * gas.R)
        zero = np.zeros_like(x)
        return np.stack([p, T, zero, zero, zero])

    a_max = np.sqrt(1.4 * 1.0 / 1.0)
    dt = 0.5 * (1.0 / n) / a_max
    return Workload("sod%d" % n, patches, RECON_UPWIND2, LIM_VANALBADA, [0.0, 0.0, 0.0],
                    [0.25, 0.5, 1.0], dt, init, steps=steps)


# ----------------------------------------------------------------------------
# Config 2: subsonic jet, 4x4x4 patches of 64^3 core cells
# ----------------------------------------------------------------------------

P_AMB = 1.0 / 1.4
T_AMB = P_AMB / (1.0 * gas.R)   # rho = 1, so a = sqrt(1.4 R T) = 1


def jet(core: int = 64, blocks: int = 4, steps: int = 1000, mach: float = 0.9, radius_cells: float = 16.0,
        noise: float = 0.01, seed: int = 1234, coflow: float = 0.1, p_jet: float = P_AMB,
        faces: str = "transmissive") -> Workload:
    """Subsonic top-hat jet entering through the z=0 face of a (core*blocks)^3 box.
    Ambient rho=1, p=1/1.4 (a=1) at rest; inflow |w| = mach inside radius_cells
    of the face centre with seeded +/-noise velocity perturbation (fixed in
    time per inflow face cell); the rest of the inflow face carries a co-flow
    of speed `coflow` (the u_far of the reference's JetFlux::xWallM,
    jetflux.h:180-185); all other domain faces transmissive; 2 seek layers on
    interior patch faces.  With coflow = 0 the rest of the face is a wall with
    zero normal velocity, and the cells beside the nozzle lip are drained to
    negative density within a few dozen steps of the impulsive start.  As in
    the jet demo (main.h:128, jetflux.h:176-178), a jet or co-flow faster than
    sound imposes p_jet inside the radius; a subsonic jet takes p from the
    boundary cell.

    faces="reference" gives the other domain faces the functors of the
    reference's JetFlux (jetflux.h:140-198, with its jet axis x mapped to z
    here): the far end (z+) a slip wall (m_WallFlux), the four lateral faces
    the far-field functor (m_FarfieldFlux, whose far state the demo leaves at
    the functor's default, the gas at rest; here the ambient state at rest)."""
    nc = core * blocks
    h = 1.0 / nc
    if faces == "transmissive":
        dom_bc = [BC_TRANSMISSIVE] * 6
    elif faces == "reference":
        dom_bc = [BC_FARFIELD] * 4 + [BC_TRANSMISSIVE, BC_SLIP]
    else:
        raise ValueError("faces must be 'transmissive' or 'reference'")
    dom_bc[4] = BC_JET
    far = _par()
    for f in range(4):
        far[f, :5] = [P_AMB, T_AMB, 0.0, 0.0, 0.0]
    patches = block_grid([nc] * 3, [blocks] * 3, [1.0] * 3, 2, [False] * 3, dom_bc,
                         domain_par=far if faces == "reference" else None)
    for P in patches:
        if P.bc[4] == BC_JET:
            par = P.bc_par[4]
            par[0] = 0.5 - P.origin[0]            # jet centre, local x
            par[1] = 0.5 - P.origin[1]            # jet centre, local y
            par[2] = radius_cells * h
            par[3] = mach * 1.0                    # a = 1
            par[4] = T_AMB
            par[5] = coflow
            par[6] = T_AMB
            par[7] = noise
            par[8] = float(seed)
            par[9] = P.gidx[0]
            par[10] = P.gidx[1]
            par[11] = p_jet
            par[12] = 1.0 if max(mach, coflow) > 1.0 else 0.0

    def init(x, y, z):
        one = np.ones_like(x)
        zero = np.zeros_like(x)
        return np.stack([P_AMB * one, T_AMB * one, zero, zero, zero])

    dt = 0.5 * h / (mach + 1.0)
    name = "jet%dx%d" % (blocks, core) + ("" if faces == "transmissive" else "ref")
    return Workload(name, patches, RECON_UPWIND2, LIM_VANALBADA, [0.0] * 3,
                    [0.25, 0.5, 1.0], dt, init, steps=steps)


# ----------------------------------------------------------------------------
# Config 5: periodic Kelvin-Helmholtz box
# ----------------------------------------------------------------------------

def kelvin_helmholtz(core: int = 64, blocks: int = 8, steps: int = 500, seed: int = 99) -> Workload:
    """Periodic shear layer: u=+0.5 for 0.25<y<0.75 (rho=2) else u=-0.5 (rho=1),
    p=1/1.4, v = 0.01 sin(4 pi x) + 0.01 U[-1,1] (seeded per global cell)."""
    nc = core * blocks
    h = 1.0 / nc
    patches = block_grid([nc] * 3, [blocks] * 3, [1.0] * 3, 2, [True] * 3, [BC_SLIP] * 6)

    def init(x, y, z):
        inside = (y > 0.25) & (y < 0.75)
        rho = np.where(inside, 2.0, 1.0)
        u = np.where(inside, 0.5, -0.5)
        # per-cell noise keyed on the global cell index (reproducible for any patching)
        gi = np.floor(x / h).astype(np.int64) % nc
        gj = np.floor(y / h).astype(np.int64) % nc
        gk = np.floor(z / h).astype(np.int64) % nc
        key = (gi * nc + gj) * nc + gk
        r = _hash_uniform(key, seed)
        v = 0.01 * np.sin(4.0 * np.pi * x) + 0.01 * r
        p = np.full_like(x, P_AMB)
        T = p / (rho * gas.R)
        return np.stack([p, T, u, v, np.zeros_like(x)])

    dt = 0.5 * h / (0.5 + np.sqrt(1.4 * P_AMB / 1.0))
    return Workload("kh%dx%d" % (blocks, core), patches, RECON_UPWIND2, LIM_VANALBADA, [1.0] * 3,
                    [0.25, 0.5, 1.0], dt, init, steps=steps)


# ----------------------------------------------------------------------------
# Viscous shear layer (SI units, so that Sutherland's viscosity matters)
# ----------------------------------------------------------------------------

def viscous_shear(core: int = 32, blocks=(2, 2, 2), cell: float = 1.0e-6, u_max: float = 30.0,
                  steps: int = 200, periodic: bool = True) -> Workload:
    """Periodic double shear layer of cubic cells of side `cell` (1 um by default)
    at 1 bar and 300 K: u = u_max tanh((y - L/4)/d) for y < L/2 and its mirror
    above, d = L/16, v = 0.05 u_max sin(2 pi x / L).  AUSM+ Upwind2 van Albada
    plus CompressibleViscFlux (the EaFlux composition of drnumCylinder/main.h:
    127-163).  The cell Reynolds number u_max h / nu is about 2 at any
    resolution, so the viscous terms are of the order of the convective ones,
    and the explicit viscous limit h^2/(6 nu) stays 8x above the convective time
    step.  periodic=False puts slip walls on the domain faces instead of
    periodic images."""
    nb = list(blocks)
    nc = [core * b for b in nb]
    h = cell
    L = [h * nc[a] for a in range(3)]
    patches = block_grid(nc, nb, L, 2, [periodic] * 3, [BC_SLIP] * 6)
    d = L[1] / 16.0

    def init(x, y, z):
        yy = np.mod(y, L[1])
        u = np.where(yy < 0.5 * L[1], u_max * np.tanh((yy - 0.25 * L[1]) / d),
                     -u_max * np.tanh((yy - 0.75 * L[1]) / d))
        v = 0.05 * u_max * np.sin(2.0 * np.pi * x / L[0])
        p = np.full_like(x, 1.0e5)
        T = np.full_like(x, 300.0)
        return np.stack([p, T, u, v, np.zeros_like(x)])

    a = np.sqrt(1.4 * gas.R * 300.0)
    dt = 0.5 * h / (u_max + a)
    return Workload("shear%dx%d" % (nb[0] * nb[1] * nb[2], core), patches, RECON_UPWIND2, LIM_VANALBADA,
                    L if periodic else [0.0] * 3, [0.25, 0.5, 1.0], dt, init, steps=steps, viscous=True)


def _hash_uniform(key: np.ndarray, seed: int) -> np.ndarray:
    """splitmix64-based uniform in [-1, 1) per integer key (vectorised)."""
    with np.errstate(over="ignore"):
        x = key.astype(np.uint64) ^ np.uint64(seed * 0x9E3779B97F4A7C15 % (1 << 64))
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    q = (x >> np.uint64(40)).astype(np.int64) - 8388608
    return q.astype(np.float64) / 8388608.0


# ----------------------------------------------------------------------------
# Config 4: level-set spheres in a 192x128x128 box
# ----------------------------------------------------------------------------

def levelset_spheres(core: int = 64, blocks=(3, 2, 2), n_spheres: int = 8, steps: int = 2000, seed: int = 7,
                     mach: float = 0.3, r_range=(6.0, 14.0)) -> Workload:
    """Freestream Mach 0.3 along x past spheres defined by the signed distance
    min_s(|x-c_s|-r_s); farfield functor on every domain face; slip level-set
    fill (3 inner layers, min relative distance 0.1) after every stage."""
    nc = [core * b for b in blocks]
    h = 1.0 / core
    L = [n * h for n in nc]
    far = _par()
    for f in range(6):
        far[f, :5] = [P_AMB, T_AMB, mach, 0.0, 0.0]
    patches = block_grid(nc, blocks, L, 2, [False] * 3, [BC_FARFIELD] * 6, domain_par=far)
    rng = np.random.default_rng(seed)
    spheres = []
    tries = 0
    while len(spheres) < n_spheres and tries < 100000:
        tries += 1
        r = rng.uniform(*r_range) * h
        c = [rng.uniform(r + 4 * h, L[a] - r - 4 * h) for a in range(3)]
        if all(np.linalg.norm(np.subtract(c, s[:3])) > r + s[3] + 2 * h for s in spheres):
            spheres.append([c[0], c[1], c[2], r])

    def init(x, y, z):
        one = np.ones_like(x)
        zero = np.zeros_like(x)
        return np.stack([P_AMB * one, T_AMB * one, mach * one, zero, zero])

    dt = 0.5 * h / (mach + 1.0)
    return Workload("ls%dspheres" % n_spheres, patches, RECON_UPWIND2, LIM_VANALBADA, [0.0] * 3,
                    [0.25, 0.5, 1.0], dt, init, spheres=np.asarray(spheres), steps=steps)


# ----------------------------------------------------------------------------
# Config 3: uneven patch batch with ratio-2 nesting
# ----------------------------------------------------------------------------

def uneven_batch(n_patches: int = 200, size_range=(16, 96), seed: int = 2024, refine_frac: float = 0.2,
                 steps: int = 100, mach: float = 0.5, pulse: float = 0.1) -> Workload:
    """Patches of uneven size, each with its own per-axis sizes drawn from
    U[size_range], on a lattice of rows: along x the patches of a row follow
    each other with a 2-4 cell overlap; along y and z each layer of rows starts
    2-4 cells below the lowest top of the layer before it, so the rows of two
    layers overlap by at least that and their patch boundaries do not line up
    (a seek layer draws on one or two donor patches; seek cells no patch
    covers keep their own update).  A fraction of the patches are ratio-2
    refined blocks nested inside a coarse parent's core.  Uniform Mach 0.5 flow
    along x plus a Gaussian density pulse (amplitude `pulse`, 0.1 by default,
    width 8 coarse cells; pulse = 0 leaves the free stream alone).  Farfield
    functors on all faces (a patch face that is also covered by a neighbour is
    overwritten by the exchange)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / 256.0
    far = _par()
    for f in range(6):
        far[f, :5] = [P_AMB, T_AMB, mach, 0.0, 0.0]
    n_coarse = int(n_patches * (1.0 - refine_frac))
    nb = 1
    while nb ** 3 < n_coarse:
        nb += 1
    slots = [(bi, bj, bk) for bk in range(nb) for bj in range(nb) for bi in range(nb)][:n_coarse]
    used = set(slots)
    sizes = {b: [int(x) for x in rng.integers(size_range[0], size_range[1] + 1, size=3)] for b in slots}
    overl = {b: [int(x) for x in rng.integers(2, 5, size=3)] for b in slots}
    # z layers, then y rows inside them, then x along each row
    g0 = {}
    zs = 0
    for bk in range(nb):
        layer = [b for b in slots if b[2] == bk]
        if not layer:
            continue
        ys = 0
        for bj in range(nb):
            row = [b for b in layer if b[1] == bj]
            if not row:
                continue
            xs = 0
            for b in sorted(row):
                g0[b] = [xs, ys, zs]
                xs += sizes[b][0] - overl[b][0]
            ys = min(g0[b][1] + sizes[b][1] for b in row) - max(overl[b][1] for b in row)
        zs = min(g0[b][2] + sizes[b][2] for b in layer) - max(overl[b][2] for b in layer)
    coarse = []
    for b in slots:
        n = sizes[b]
        seek = []
        for a in range(3):
            lower = tuple(b[q] - (q == a) for q in range(3))
            upper = tuple(b[q] + (q == a) for q in range(3))
            seek += [2 if lower in used else 0, 2 if upper in used else 0]
        org = [g0[b][a] * h for a in range(3)]
        L = [n[a] * h for a in range(3)]
        coarse.append(Patch(n[0], n[1], n[2], seek, org, L, [BC_FARFIELD] * 6, far.copy(), list(g0[b])))
    patches: List[Patch] = list(coarse)
    n_fine = n_patches - len(coarse)
    for f in range(n_fine):
        parent = coarse[int(rng.integers(0, len(coarse)))]
        # nested block of the parent's core, refined by 2, 2 seek layers all round
        pc = [[parent.seek[2 * a], [parent.ni, parent.nj, parent.nk][a] - parent.seek[2 * a + 1]] for a in range(3)]
        ext = [max(4, int(rng.integers(4, max(5, (pc[a][1] - pc[a][0]) // 2)))) for a in range(3)]
        lo = [pc[a][0] + 2 + int(rng.integers(0, max(1, pc[a][1] - pc[a][0] - ext[a] - 3))) for a in range(3)]
        n = [2 * ext[a] + 4 for a in range(3)]
        hf = h / 2.0
        org = [parent.origin[a] + lo[a] * h - 2 * hf for a in range(3)]
        L = [n[a] * hf for a in range(3)]
        patches.append(Patch(n[0], n[1], n[2], [2] * 6, org, L, [BC_FARFIELD] * 6, far.copy(),
                             [0, 0, 0]))
    ext_hi = [max(p.origin[a] + p.length[a] for p in patches) for a in range(3)]
    centre = [0.5 * ext_hi[a] for a in range(3)]
    width = 8 * h

    def init(x, y, z):
        r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
        rho = 1.0 + pulse * np.exp(-r2 / (2 * width * width))
        p = np.full_like(x, P_AMB)
        T = p / (rho * gas.R)
        zero = np.zeros_like(x)
        return np.stack([p, T, mach * np.ones_like(x), zero, zero])

    dt = 0.5 * (h / 2.0) / (mach + 1.2)
    return Workload("uneven%d" % n_patches, patches, RECON_UPWIND2, LIM_VANALBADA, [0.0] * 3,
                    [0.25, 0.5, 1.0], dt, init, steps=steps)


# ----------------------------------------------------------------------------
# Rotated patches: a star of rotated patches around an unrotated centre patch
# ----------------------------------------------------------------------------

def _rot_z(phi: float) -> np.ndarray:
    c, s = np.cos(phi), np.sin(phi)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def rotated_star(core: int = 16, seek: int = 2, steps: int = 10, flow=(0.3, 0.2, 0.1), pulse: float = 0.1,
                 tilt_deg: float = 0.0, arms: str = "xyz") -> Workload:
    """An unrotated centre patch C at the inertial origin and three rotated
    patches on its +x, +y and +z sides, each overlapping C by `seek` seek
    layers on both sides (patch files give each patch's base vectors,
    patch.cpp:63-85):

      C   cells [0, core+seek)^3, seek layers on its x, y and z max faces;
      RX  x in [core-seek, 2 core), y, z in [0, core): turned 90 degrees about z
          (local i = +y, local j = -x, local k = +z), seek layers on local j max;
      RY  y in [core-seek, 2 core), x, z in [0, core): turned 90 degrees about x
          (local i = +x, local j = +z, local k = -y), seek layers on local k max;
      RZ  z in [core-seek, 2 core), x, y in [0, core): turned 180 degrees about z
          (local i = -x, local j = -y, local k = +z), seek layers on local k min.

    Every exchange relation joins a rotated patch with C, which is the
    identity transform: for such pairs the reference's relative transform
    (CoordTransformVV::setTransFromTo) is the geometric one.  With tilt_deg = 0
    the rotated patches' cells coincide with C's, so every donor weight is
    exactly 1 and the momentum is turned by an exact signed permutation.
    tilt_deg turns RX further about the z axis through the middle of its
    overlap with C (trilinear donor weights, a general turn matrix); the
    tilted RX then also overlaps RY and RZ, two turned patches, for which
    the reference's composition is not the geometric transform (unless both
    turn about one axis).  `arms` selects which of RX, RY, RZ are present
    (C's faces without an arm keep no seek layers).

    Uniform flow `flow` (times the sound speed, inertial frame) plus a
    Gaussian density pulse (amplitude `pulse`) in C; transmissive faces."""
    h = 1.0 / 64.0
    c, s = core, seek
    tr = [BC_TRANSMISSIVE] * 6

    def patch(n, seeks, origin, base_i, base_j):
        return Patch(n[0], n[1], n[2], list(seeks), [float(o) for o in origin], [n[a] * h for a in range(3)],
                     list(tr), _par(), [0, 0, 0], [float(b) for b in base_i], [float(b) for b in base_j])

    sx, sy, sz = (s if a in arms else 0 for a in "xyz")
    C = patch((c + sx, c + sy, c + sz), (0, sx, 0, sy, 0, sz), (0.0, 0.0, 0.0), (1, 0, 0), (0, 1, 0))
    rx_o = np.array([2 * c * h, 0.0, 0.0])
    rx_bi, rx_bj = np.array([0.0, 1.0, 0.0]), np.array([-1.0, 0.0, 0.0])
    if tilt_deg != 0.0:
        R = _rot_z(np.radians(tilt_deg))
        pivot = np.array([c * h, 0.5 * c * h, 0.0])
        rx_o = pivot + R @ (rx_o - pivot)
        rx_bi, rx_bj = R @ rx_bi, R @ rx_bj
    RX = patch((c, c + s, c), (0, 0, 0, s, 0, 0), rx_o, rx_bi, rx_bj)
    RY = patch((c, c, c + s), (0, 0, 0, 0, 0, s), (0.0, 2 * c * h, 0.0), (1, 0, 0), (0, 0, 1))
    RZ = patch((c, c, c + s), (0, 0, 0, 0, s, 0), (c * h, c * h, (c - s) * h), (-1, 0, 0), (0, -1, 0))
    patches = [C] + [P for a, P in zip("xyz", (RX, RY, RZ)) if a in arms]
    centre = np.array([0.5 * c * h] * 3)
    width = 3 * h
    a0 = 1.0

    def init(x, y, z):
        r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
        rho = 1.0 + pulse * np.exp(-r2 / (2 * width * width))
        p = np.full_like(x, P_AMB)
        T = p / (rho * gas.R)
        return np.stack([p, T, flow[0] * a0 * np.ones_like(x), flow[1] * a0 * np.ones_like(x),
                         flow[2] * a0 * np.ones_like(x)])

    dt = 0.4 * h / (np.linalg.norm(flow) * a0 + 1.2 * a0)
    return Workload("rotstar%d" % core + ("_tilt%g" % tilt_deg if tilt_deg else "") +
                    ("" if arms == "xyz" else "_" + arms), patches, RECON_UPWIND2,
                    LIM_VANALBADA, [0.0] * 3, [0.25, 0.5, 1.0], dt, init, steps=steps)
