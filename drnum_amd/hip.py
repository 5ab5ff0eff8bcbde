"""ctypes binding of libdrnhip.so (include/drnhip.h).

The library is built in-tree (drnum_amd/lib/libdrnhip.so) by
``python -m drnum_amd.build``.  There is no fallback: if the library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdrnhip.so")

DRN_F32, DRN_F64 = 0, 1
DRN_LAUNCH_BATCHED, DRN_LAUNCH_PER_PATCH = 0, 1
DRN_ARITH_IEEE, DRN_ARITH_FAST = 0, 1
DRN_LS_SLIP, DRN_LS_SWALL = 0, 1
DRN_CELL_INACTIVE, DRN_CELL_INSIDE = 1, 2
NPAR = 16


class PatchDesc(C.Structure):
    _fields_ = [
        ("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32),
        ("seek", C.c_int32 * 6),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("bc", C.c_int32 * 6),
        ("bc_par", (C.c_double * NPAR) * 6),
    ]


_SIGS = {
    "drn_ctx_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "drn_ctx_destroy": (None, [C.c_void_p]),
    "drn_last_error": (C.c_char_p, [C.c_void_p]),
    "drn_precision": (C.c_int, [C.c_void_p]),
    "drn_set_operator": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "drn_set_viscous": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_patch_add": (C.c_int, [C.c_void_p, C.POINTER(PatchDesc), C.POINTER(C.c_int)]),
    "drn_layout": (C.c_int, [C.c_void_p]),
    "drn_num_patches": (C.c_int, [C.c_void_p]),
    "drn_patch_cells": (C.c_size_t, [C.c_void_p, C.c_int]),
    "drn_patch_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "drn_patch_download": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "drn_copy_field": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "drn_swap_fields": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "drn_stage": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]),
    "drn_exchange_set": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                   C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "drn_exchange_set_turn": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "drn_set_cell_flags": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "drn_exchange_commit": (C.c_int, [C.c_void_p]),
    "drn_exchange": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_set_exchange_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_exchange_fused_cells": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "drn_levelset_set": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "drn_levelset_commit": (C.c_int, [C.c_void_p]),
    "drn_levelset_fill": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "drn_set_levelset_op": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_variable_difference": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p]),
    "drn_rk_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double), C.c_int]),
    "drn_rk_graph_run": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double), C.c_int]),
    "drn_core_sums": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "drn_core_sums_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "drn_diagnostics": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "drn_cfl_axes": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "drn_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "drn_set_tiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "drn_get_tiling": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_num_tilings": (C.c_int, []),
    "drn_post_variables": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p]),
    "drn_set_arith": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_timing": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "drn_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "drn_sync": (C.c_int, [C.c_void_p]),
}

_lib = None


def lib():
    """Load libdrnhip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdrnhip.so is not built (%s); run python -m drnum_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


class DrnError(RuntimeError):
    pass


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context of libdrnhip (precision np.float32 or np.float64)."""

    def __init__(self, dtype=np.float64, device: int = 0):
        self.dtype = np.dtype(dtype)
        self.L = lib()
        h = C.c_void_p()
        prec = DRN_F32 if self.dtype == np.float32 else DRN_F64
        rc = self.L.drn_ctx_create(device, prec, C.byref(h))
        if rc != 0:
            raise DrnError("drn_ctx_create failed (%d): no usable HIP device %d" % (rc, device))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.drn_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.drn_last_error(self.h)
            raise DrnError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # ---- set-up -----------------------------------------------------------
    def set_operator(self, recon, limiter, viscous=False):
        self._chk(self.L.drn_set_operator(self.h, recon, limiter), "drn_set_operator")
        self._chk(self.L.drn_set_viscous(self.h, 1 if viscous else 0), "drn_set_viscous")

    def add_patch(self, ni, nj, nk, seek, dxyz, bc, bc_par):
        d = PatchDesc()
        d.ni, d.nj, d.nk = ni, nj, nk
        for a in range(6):
            d.seek[a] = int(seek[a])
            d.bc[a] = int(bc[a])
            for q in range(NPAR):
                d.bc_par[a][q] = float(bc_par[a][q])
        d.dx, d.dy, d.dz = (float(x) for x in dxyz)
        pid = C.c_int()
        self._chk(self.L.drn_patch_add(self.h, C.byref(d), C.byref(pid)), "drn_patch_add")
        return pid.value

    def layout(self):
        self._chk(self.L.drn_layout(self.h), "drn_layout")

    def patch_cells(self, pid) -> int:
        return int(self.L.drn_patch_cells(self.h, pid))

    def upload(self, pid, field, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=self.dtype).reshape(-1)
        if a.size != 5 * self.patch_cells(pid):
            raise DrnError("upload: patch %d holds %d cells, got %d values" % (pid, self.patch_cells(pid), a.size))
        self._chk(self.L.drn_patch_upload(self.h, pid, field, _p(a)), "drn_patch_upload")

    def download(self, pid, field=0) -> np.ndarray:
        a = np.empty(5 * self.patch_cells(pid), dtype=self.dtype)
        self._chk(self.L.drn_patch_download(self.h, pid, field, _p(a)), "drn_patch_download")
        return a.reshape(5, -1)

    def set_exchange(self, pid, tables):
        """tables: dict(receivers, donors=[(patch, nrec, stride, rec_start, wi_start)], rec_concat,
        donor_index, donor_weight) in the reference's direct transfer-list form."""
        donors = tables["donors"]
        dp = np.array([d[0] for d in donors] or [0], dtype=np.int32)
        info = np.array([[d[1], d[2], d[3], d[4]] for d in donors] or [[0, 0, 0, 0]], dtype=np.int64)
        ru = np.ascontiguousarray(tables["receivers"], dtype=np.int64)
        rc_ = np.ascontiguousarray(tables["rec_concat"], dtype=np.int64)
        di = np.ascontiguousarray(tables["donor_index"], dtype=np.int64)
        dw = np.ascontiguousarray(tables["donor_weight"], dtype=np.float64)
        keep = (ru, rc_, di, dw, dp, info)
        self._chk(self.L.drn_exchange_set(self.h, pid, len(donors), _p(dp), _p(info), ru.size, _p(ru), rc_.size,
                                          _p(rc_), di.size, _p(di), _p(dw)), "drn_exchange_set")
        del keep
        if tables.get("turn") is not None and len(donors) > 0:
            t = np.ascontiguousarray(tables["turn"], dtype=np.float64).reshape(len(donors), 9)
            self._chk(self.L.drn_exchange_set_turn(self.h, pid, len(donors), _p(t)), "drn_exchange_set_turn")

    def set_cell_flags(self, pid, flags):
        """per-cell flags of patch pid (DRN_CELL_INACTIVE = 1, DRN_CELL_INSIDE = 2), None clears"""
        if flags is None:
            self._chk(self.L.drn_set_cell_flags(self.h, pid, None), "drn_set_cell_flags")
            return
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        self._chk(self.L.drn_set_cell_flags(self.h, pid, _p(f)), "drn_set_cell_flags")

    def commit_exchange(self):
        self._chk(self.L.drn_exchange_commit(self.h), "drn_exchange_commit")

    def set_levelset(self, pid, cells, normals, donors, weights):
        c = np.ascontiguousarray(cells, dtype=np.int64)
        n = np.ascontiguousarray(normals, dtype=np.float64)
        d = np.ascontiguousarray(donors, dtype=np.int64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        self._chk(self.L.drn_levelset_set(self.h, pid, c.size, _p(c), _p(n), _p(d), _p(w)), "drn_levelset_set")

    def commit_levelset(self):
        self._chk(self.L.drn_levelset_commit(self.h), "drn_levelset_commit")

    # ---- compute ----------------------------------------------------------
    def stage(self, src, base, dst, factor, launch=DRN_LAUNCH_BATCHED):
        self._chk(self.L.drn_stage(self.h, src, base, dst, float(factor), launch), "drn_stage")

    def swap_fields(self, a, b):
        self._chk(self.L.drn_swap_fields(self.h, a, b), "drn_swap_fields")

    def copy_field(self, src, dst):
        self._chk(self.L.drn_copy_field(self.h, src, dst), "drn_copy_field")

    def exchange(self, field=0):
        self._chk(self.L.drn_exchange(self.h, field), "drn_exchange")

    def set_exchange_fusion(self, on=True):
        """the cell stage kernels fill the aligned z-face receivers inside rk_step (default on)"""
        self._chk(self.L.drn_set_exchange_fusion(self.h, int(bool(on))), "drn_set_exchange_fusion")

    def exchange_fused_cells(self, launch=DRN_LAUNCH_BATCHED) -> int:
        """receivers the stage kernel fills itself in rk_step for this launch form"""
        n = C.c_int64(0)
        self._chk(self.L.drn_exchange_fused_cells(self.h, launch, C.byref(n)), "drn_exchange_fused_cells")
        return int(n.value)

    def levelset_fill(self, field=0, relax=0.5):
        self._chk(self.L.drn_levelset_fill(self.h, field, relax), "drn_levelset_fill")

    def variable_difference(self, patch, field1, var1, field2, var2):
        """Patch::computeVariableDifference on the device: (max_norm, l2_norm)"""
        mx, l2 = C.c_double(0), C.c_double(0)
        self._chk(self.L.drn_variable_difference(self.h, patch, field1, var1, field2, var2, C.byref(mx), C.byref(l2)),
                  "drn_variable_difference")
        return mx.value, l2.value

    def set_levelset_op(self, op):
        """DRN_LS_SLIP (LSOBCCompressibleEulerOp, default) or DRN_LS_SWALL (LSOBCCompressibleSWallOp)"""
        self._chk(self.L.drn_set_levelset_op(self.h, op), "drn_set_levelset_op")

    def rk_step(self, dt, alpha, launch=DRN_LAUNCH_BATCHED):
        a = (C.c_double * len(alpha))(*alpha)
        self._chk(self.L.drn_rk_step(self.h, float(dt), len(alpha), a, launch), "drn_rk_step")

    def rk_graph_run(self, dt, alpha, n_steps):
        a = (C.c_double * len(alpha))(*alpha)
        self._chk(self.L.drn_rk_graph_run(self.h, float(dt), len(alpha), a, n_steps), "drn_rk_graph_run")

    def core_sums(self, field=0):
        out = (C.c_double * 5)()
        self._chk(self.L.drn_core_sums(self.h, field, out), "drn_core_sums")
        return np.array(out[:])

    def core_sums_noalloc(self, d_out, field=0):
        """drn_core_sums_noalloc: the sums into 5 doubles of device memory at d_out
        (an address), asynchronous on the context stream"""
        self._chk(self.L.drn_core_sums_noalloc(self.h, field, C.c_void_p(d_out)), "drn_core_sums_noalloc")

    def diagnostics(self, field=0, dt=0.0):
        out = (C.c_double * 4)()
        self._chk(self.L.drn_diagnostics(self.h, field, float(dt), out), "drn_diagnostics")
        return dict(cfl=out[0], rho_min=out[1], rho_max=out[2], nonfinite=out[3])

    def cfl_axes(self, field=0, dt=0.0):
        """the CFL number over all three axes (no counterpart in the reference)"""
        out = C.c_double()
        self._chk(self.L.drn_cfl_axes(self.h, field, float(dt), C.byref(out)), "drn_cfl_axes")
        return out.value

    def autotune(self, launch=DRN_LAUNCH_BATCHED):
        c = C.c_int()
        self._chk(self.L.drn_autotune(self.h, launch, C.byref(c)), "drn_autotune")
        return c.value

    def set_tiling(self, choice, launch=DRN_LAUNCH_BATCHED):
        self._chk(self.L.drn_set_tiling(self.h, launch, choice), "drn_set_tiling")

    def get_tiling(self, launch=DRN_LAUNCH_BATCHED):
        return self.L.drn_get_tiling(self.h, launch)

    POST_NAMES = ("Ma", "p", "T", "rho", "S", "u", "v", "w")

    def post_variables(self, patch, names=POST_NAMES, field=0, ref_p=1e5, ref_T=300.0):
        """Post-processing variables of one patch computed on the device:
        dict name -> (N,) array in the context precision."""
        mask = 0
        for nm in names:
            mask |= 1 << self.POST_NAMES.index(nm)
        sel = [nm for nm in self.POST_NAMES if mask & (1 << self.POST_NAMES.index(nm))]
        n = self.patch_cells(patch)
        out = np.zeros((len(sel), n), dtype=self.dtype)
        self._chk(self.L.drn_post_variables(self.h, field, patch, mask, float(ref_p), float(ref_T),
                                            out.ctypes.data_as(C.c_void_p)), "drn_post_variables")
        return dict(zip(sel, out))

    def num_tilings(self):
        return self.L.drn_num_tilings()

    def set_arith(self, mode):
        self._chk(self.L.drn_set_arith(self.h, mode), "drn_set_arith")

    def timing_enable(self, on=True):
        self._chk(self.L.drn_timing_enable(self.h, int(on)), "drn_timing_enable")

    def timing(self, which=0):
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self.L.drn_timing(self.h, which, C.byref(ms), C.byref(n)), "drn_timing")
        return ms.value, n.value

    def sync(self):
        self._chk(self.L.drn_sync(self.h), "drn_sync")
