"""Set up a libdrnhip context for a workload (configs.Workload).

This is the Python-side convenience driver used by bench.py and the tests; the
C++ drop-in (HIP_CartesianIterator + RungeKutta in drnum_amd/host) drives the
same C ABI.
"""
from __future__ import annotations

import numpy as np

from . import configs, gas, hip


def patch_spacing(P: configs.Patch, dtype) -> list:
    """CartesianPatch::computeDeltas (cartesianpatch.cpp:164-168): m_Dx = m_Lx/m_NumI in `real`."""
    f = np.dtype(dtype).type
    n = (P.ni, P.nj, P.nk)
    return [float(f(P.length[a]) / f(n[a])) for a in range(3)]


def device_bc_par(P: configs.Patch, dtype) -> np.ndarray:
    """Face parameters as the kernels expect them: the farfield state is handed
    over in conserved variables (CompressibleFarfieldFlux::setFarfield,
    compressiblefarfieldflux.h:74-77, stores m_Var = primitiveToConservative(p,T,u,v,w))."""
    par = np.array(P.bc_par, dtype=np.float64, copy=True)
    for f in range(6):
        if P.bc[f] == configs.BC_FARFIELD:
            cons = gas.prim_to_cons(*[np.asarray(par[f, q]) for q in range(5)], dtype=dtype)
            par[f, :5] = cons.astype(np.float64)
    return par


class GpuGrid:
    """A workload laid out on one device in one precision."""

    def __init__(self, workload: configs.Workload, dtype, device: int = 0):
        self.w = workload
        self.dtype = np.dtype(dtype)
        self.ctx = hip.Context(self.dtype, device)
        self.ctx.set_operator(workload.recon, workload.limiter, workload.viscous)
        for P in workload.patches:
            self.ctx.add_patch(P.ni, P.nj, P.nk, P.seek, patch_spacing(P, self.dtype), P.bc,
                               device_bc_par(P, self.dtype))
        self.ctx.layout()

    def upload_initial(self):
        for p in range(len(self.w.patches)):
            self.ctx.upload(p, 0, self.w.initial_state(p, self.dtype))

    def upload_state(self, states, field=0):
        for p, U in enumerate(states):
            self.ctx.upload(p, field, U)

    def set_tables(self, tables):
        """tables[p]: reference-form transfer lists of receiver patch p (or None)."""
        for p, t in enumerate(tables):
            if t is not None and len(t["donors"]) > 0:
                self.ctx.set_exchange(p, t)
        self.ctx.commit_exchange()

    def set_levelset(self, ls):
        """ls[p]: dict(cells, normals (n,3), donors (n,8), weights (n,8)) in flat fill order."""
        for p, t in enumerate(ls):
            if t is not None and len(t["cells"]) > 0:
                self.ctx.set_levelset(p, t["cells"], t["normals"], t["donors"], t["weights"])
        self.ctx.commit_levelset()

    def state(self, field=0):
        return [self.ctx.download(p, field) for p in range(len(self.w.patches))]

    def rk_step(self, launch=hip.DRN_LAUNCH_BATCHED):
        self.ctx.rk_step(self.w.dt, self.w.alpha, launch)

    def close(self):
        self.ctx.close()
