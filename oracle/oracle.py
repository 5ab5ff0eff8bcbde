"""ctypes front end of the C oracle (oracle/liboracle.so).  Test infrastructure.

    g = OracleGrid(workload, np.float64)      # o64_* (FP64 build of the reference)
    g.init_state()                             # field 0 <- workload initial state
    g.compute_dependencies()
    g.rk_step(dt, alpha)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NPAR = 16


class PatchDesc(C.Structure):
    _fields_ = [
        ("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32),
        ("seek", C.c_int32 * 6),
        ("origin", C.c_double * 3),
        ("length", C.c_double * 3),
        ("bc", C.c_int32 * 6),
        ("bc_par", (C.c_double * NPAR) * 6),
        ("base_i", C.c_double * 3),
        ("base_j", C.c_double * 3),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for pfx, rt in (("o32_", C.c_float), ("o64_", C.c_double)):
            rp = C.POINTER(rt)
            sig = {
                "prim_to_cons": (None, [rt, rt, rt, rt, rt, rp]),
                "limiter": (rt, [C.c_int, rt, rt]),
                "ausm_flux": (None, [rp, rp, C.c_int, rt, rp]),
                "gas_mu": (rt, [rp]),
                "visc_face": (None, [rp, rp, rp, rp, rp, rp, rt, C.c_int, rt, rp]),
                "grid_set_viscous": (None, [C.c_void_p, C.c_int]),
                "grid_create": (C.c_void_p, [C.c_int, C.POINTER(PatchDesc), C.c_int, C.c_int]),
                "grid_destroy": (None, [C.c_void_p]),
                "grid_patch_cells": (C.c_size_t, [C.c_void_p, C.c_int]),
                "grid_field": (rp, [C.c_void_p, C.c_int, C.c_int]),
                "grid_dx": (rt, [C.c_void_p, C.c_int, C.c_int]),
                "grid_set_const": (None, [C.c_void_p, C.c_int, rp]),
                "grid_compute": (None, [C.c_void_p, rt]),
                "grid_residual": (None, [C.c_void_p, C.c_int, rp]),
                "grid_copy_field": (None, [C.c_void_p, C.c_int, C.c_int]),
                "grid_compute_dependencies": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
                "grid_num_receivers": (C.c_int, [C.c_void_p, C.c_int]),
                "grid_num_no_donor": (C.c_int, [C.c_void_p, C.c_int]),
                "grid_receivers": (C.POINTER(C.c_int64), [C.c_void_p, C.c_int]),
                "grid_num_donors": (C.c_int, [C.c_void_p, C.c_int]),
                "grid_donor_info": (None, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
                "grid_num_concat": (C.c_int64, [C.c_void_p, C.c_int, C.c_int]),
                "grid_rec_concat": (C.POINTER(C.c_int64), [C.c_void_p, C.c_int]),
                "grid_donor_index": (C.POINTER(C.c_int64), [C.c_void_p, C.c_int]),
                "grid_donor_weight": (rp, [C.c_void_p, C.c_int]),
                "grid_exchange": (None, [C.c_void_p, C.c_int]),
                "grid_set_cell_flags": (None, [C.c_void_p, C.c_int, C.c_void_p]),
                "grid_levelset": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int,
                                            C.c_double]),
                "grid_levelset_g": (rp, [C.c_void_p, C.c_int]),
                "grid_ls_num_cells": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
                "grid_ls_cells": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), rp, rp]),
                "grid_levelset_fill": (None, [C.c_void_p, C.c_int, C.c_int]),
                "grid_set_levelset_op": (None, [C.c_void_p, C.c_int]),
                "grid_variable_difference": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                    C.c_void_p, C.c_void_p]),
                "grid_rk_step": (None, [C.c_void_p, rt, C.c_int, rp]),
                "grid_set_threads": (None, [C.c_void_p, C.c_int]),
                "grid_core_sums": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
                "grid_donor_matrix": (None, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
                "grid_patch_transform": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, pfx + name)
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib


def _pfx(dtype):
    return "o32_" if np.dtype(dtype) == np.float32 else "o64_"


def _ctype(dtype):
    return C.c_float if np.dtype(dtype) == np.float32 else C.c_double


def fn(dtype, name):
    return getattr(lib(), _pfx(dtype) + name)


def descs(patches):
    arr = (PatchDesc * len(patches))()
    for d, P in zip(arr, patches):
        d.ni, d.nj, d.nk = P.ni, P.nj, P.nk
        for a in range(6):
            d.seek[a] = P.seek[a]
            d.bc[a] = P.bc[a]
            for q in range(NPAR):
                d.bc_par[a][q] = float(P.bc_par[a][q])
        bi = getattr(P, "base_i", (1.0, 0.0, 0.0))
        bj = getattr(P, "base_j", (0.0, 1.0, 0.0))
        for a in range(3):
            d.origin[a] = P.origin[a]
            d.length[a] = P.length[a]
            d.base_i[a] = float(bi[a])
            d.base_j[a] = float(bj[a])
    return arr


def prim_to_cons(p, T, u, v, w, dtype):
    rt = _ctype(dtype)
    out = (rt * 5)()
    fn(dtype, "prim_to_cons")(rt(p), rt(T), rt(u), rt(v), rt(w), out)
    return np.array(out[:], dtype=dtype)


def limiter(lim, d1, d2, dtype):
    rt = _ctype(dtype)
    return fn(dtype, "limiter")(lim, rt(d1), rt(d2))


def ausm_flux(var_l, var_r, direction, area, dtype):
    rt = _ctype(dtype)
    vl = np.ascontiguousarray(var_l, dtype=dtype)
    vr = np.ascontiguousarray(var_r, dtype=dtype)
    out = np.zeros(5, dtype=dtype)
    p = C.POINTER(rt)
    fn(dtype, "ausm_flux")(vl.ctypes.data_as(p), vr.ctypes.data_as(p), direction, rt(area),
                           out.ctypes.data_as(p))
    return out


def gas_mu(var, dtype):
    """PerfectGas::mu (perfectgas.h:100-111) of a conserved state."""
    v = np.ascontiguousarray(var, dtype=dtype)
    return fn(dtype, "gas_mu")(v.ctypes.data_as(C.POINTER(_ctype(dtype))))


def visc_face(var_l, var_r, gl_a, gr_a, gl_b, gr_b, D, direction, area, dtype, flux=None):
    """CompressibleViscFlux::{x,y,z}Field (compressibleviscflux.h:33-268) on given
    states and cell gradients; returns flux - viscous terms (flux defaults to 0)."""
    rt = _ctype(dtype)
    p = C.POINTER(rt)
    arrs = [np.ascontiguousarray(a, dtype=dtype) for a in (var_l, var_r, gl_a, gr_a, gl_b, gr_b)]
    out = np.zeros(5, dtype=dtype) if flux is None else np.array(flux, dtype=dtype)
    fn(dtype, "visc_face")(*[a.ctypes.data_as(p) for a in arrs], rt(D), direction, rt(area), out.ctypes.data_as(p))
    return out


class OracleGrid:
    """A grid of patches held by the oracle in one precision (np.float32 or np.float64)."""

    def __init__(self, workload, dtype, threads: int = 1):
        self.w = workload
        self.dtype = np.dtype(dtype)
        self.rt = _ctype(dtype)
        self._d = descs(workload.patches)
        self.h = fn(dtype, "grid_create")(len(workload.patches), self._d, workload.recon, workload.limiter)
        self.np = len(workload.patches)
        fn(dtype, "grid_set_threads")(self.h, threads)
        if getattr(workload, "viscous", False):
            fn(dtype, "grid_set_viscous")(self.h, 1)

    def __del__(self):
        try:
            if self.h:
                fn(self.dtype, "grid_destroy")(self.h)
                self.h = None
        except Exception:
            pass

    def _f(self, name):
        return fn(self.dtype, name)

    def cells(self, p):
        return int(self._f("grid_patch_cells")(self.h, p))

    def field(self, p, field=0) -> np.ndarray:
        """(5, N) view of a field of patch p (writable, aliases oracle memory)."""
        n = self.cells(p)
        ptr = self._f("grid_field")(self.h, p, field)
        return np.ctypeslib.as_array(ptr, shape=(5 * n,)).reshape(5, n)

    def init_state(self):
        for p in range(self.np):
            self.field(p, 0)[:] = self.w.initial_state(p, self.dtype)

    def set_threads(self, n):
        self._f("grid_set_threads")(self.h, n)

    def compute(self, factor):
        self._f("grid_compute")(self.h, self.rt(factor))

    def residual(self, p):
        out = np.zeros(5 * self.cells(p), dtype=self.dtype)
        self._f("grid_residual")(self.h, p, out.ctypes.data_as(C.POINTER(self.rt)))
        return out.reshape(5, -1)

    def copy_field(self, src, dst):
        self._f("grid_copy_field")(self.h, src, dst)

    def compute_dependencies(self):
        per = (C.c_double * 3)(*self.w.periodic)
        r = self._f("grid_compute_dependencies")(self.h, per)
        if r != 0:
            raise RuntimeError("oracle: donor search hit a reference BUG condition")

    def exchange(self, field=0):
        self._f("grid_exchange")(self.h, field)

    def set_cell_flags(self, p, flags):
        """per-cell flags of patch p (1 inactive, 2 inside); None clears"""
        if flags is None:
            self._f("grid_set_cell_flags")(self.h, p, None)
            return
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.size == self.cells(p)
        self._f("grid_set_cell_flags")(self.h, p, f.ctypes.data_as(C.c_void_p))

    def rk_step(self, dt=None, alpha=None):
        alpha = self.w.alpha if alpha is None else alpha
        dt = self.w.dt if dt is None else dt
        a = (self.rt * len(alpha))(*alpha)
        self._f("grid_rk_step")(self.h, self.rt(dt), len(alpha), a)

    def core_sums(self, field=0):
        out = (C.c_double * 5)()
        self._f("grid_core_sums")(self.h, field, out)
        return np.array(out[:])

    # ---- donor tables -------------------------------------------------------
    def tables(self, p):
        """Direct transfer lists of receiver patch p (patch.cpp:128-218)."""
        f = self._f
        def arr(ptr, n, dtype):
            if n == 0 or not ptr:
                return np.zeros(0, dtype=dtype)
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

        nrec = f("grid_num_receivers")(self.h, p)
        recv = arr(f("grid_receivers")(self.h, p), nrec, np.int64)
        nd = f("grid_num_donors")(self.h, p)
        donors = []
        info = (C.c_int64 * 5)()
        for d in range(nd):
            f("grid_donor_info")(self.h, p, d, info)
            donors.append(tuple(int(x) for x in info))
        nrc = int(f("grid_num_concat")(self.h, p, 0))
        nwi = int(f("grid_num_concat")(self.h, p, 1))
        rc = arr(f("grid_rec_concat")(self.h, p), nrc, np.int64)
        di = arr(f("grid_donor_index")(self.h, p), nwi, np.int64)
        dw = arr(f("grid_donor_weight")(self.h, p), nwi, self.dtype)
        turn = np.array([self.donor_matrix(p, d).ravel() for d in range(nd)]).reshape(nd, 9)
        return dict(receivers=recv, donors=donors, rec_concat=rc, donor_index=di, donor_weight=dw,
                    no_donor=int(f("grid_num_no_donor")(self.h, p)), turn=turn)

    def donor_matrix(self, p, d):
        """the turn of vector variables of donor d of receiver patch p (3x3)"""
        out = (C.c_double * 9)()
        self._f("grid_donor_matrix")(self.h, p, d, out)
        return np.array(out[:]).reshape(3, 3)

    def patch_transform(self, p):
        """patch p's this->inertial transform as Patch::xyzoCell applies it: (A, b)"""
        out = (C.c_double * 12)()
        self._f("grid_patch_transform")(self.h, p, out)
        a = np.array(out[:])
        return a[:9].reshape(3, 3), a[9:]

    # ---- level set ----------------------------------------------------------
    def levelset(self, spheres, n_inner=3, n_outer=0, min_inner_rel=0.1):
        s = np.ascontiguousarray(spheres, dtype=np.float64).ravel()
        r = self._f("grid_levelset")(self.h, len(s) // 4, s.ctypes.data_as(C.POINTER(C.c_double)), n_inner,
                                     n_outer, min_inner_rel)
        if r != 0:
            raise RuntimeError("oracle: level-set mirror search hit a reference BUG condition")

    def levelset_g(self, p):
        return np.ctypeslib.as_array(self._f("grid_levelset_g")(self.h, p), shape=(self.cells(p),)).copy()

    def ls_cells(self, p, layer, outer=False):
        n = self._f("grid_ls_num_cells")(self.h, p, layer, int(outer))
        cells = np.zeros(max(n, 1), np.int64)
        donors = np.zeros((max(n, 1), 8), np.int64)
        g = np.zeros((max(n, 1), 4), self.dtype)
        w = np.zeros((max(n, 1), 8), self.dtype)
        P64 = C.POINTER(C.c_int64)
        PR = C.POINTER(self.rt)
        self._f("grid_ls_cells")(self.h, p, layer, int(outer), cells.ctypes.data_as(P64), donors.ctypes.data_as(P64),
                                 g.ctypes.data_as(PR), w.ctypes.data_as(PR))
        return dict(cells=cells[:n], donors=donors[:n], g=g[:n], weights=w[:n])

    def levelset_fill(self, field=0, abuse_field=2):
        self._f("grid_levelset_fill")(self.h, field, abuse_field)

    def variable_difference(self, p, f1, v1, f2, v2):
        """Patch::computeVariableDifference: (max_norm, l2_norm) in the state type"""
        mx, l2 = C.c_double(0), C.c_double(0)
        self._f("grid_variable_difference")(self.h, p, f1, v1, f2, v2, C.byref(mx), C.byref(l2))
        return mx.value, l2.value

    def set_levelset_op(self, op):
        """0: LSOBCCompressibleEulerOp (slip, the default), 1: LSOBCCompressibleSWallOp"""
        self._f("grid_set_levelset_op")(self.h, op)
