"""ctypes binding of libdrnum.so (include/drnum_host.h): the host-side patch grid,
donor tables and level-set tables of the drop-in, built in FP64 geometry."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import configs, hip

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdrnum.so")

_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_SIGS = {
    "drnh_grid_create": (C.c_void_p, [C.c_int]),
    "drnh_grid_destroy": (None, [C.c_void_p]),
    "drnh_last_error": (C.c_char_p, [C.c_void_p]),
    "drnh_grid_add_patch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32), _DP, _DP]),
    "drnh_grid_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "drnh_grid_num_patches": (C.c_int, [C.c_void_p]),
    "drnh_grid_patch_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), _DP]),
    "drnh_grid_set_faces": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), _DP]),
    "drnh_grid_compute_dependencies": (C.c_int, [C.c_void_p, _DP, C.c_int]),
    "drnh_grid_num_receivers": (C.c_int64, [C.c_void_p, C.c_int]),
    "drnh_grid_num_no_donor": (C.c_int64, [C.c_void_p, C.c_int]),
    "drnh_grid_receivers": (C.c_int, [C.c_void_p, C.c_int, _I64P]),
    "drnh_grid_num_donors": (C.c_int, [C.c_void_p, C.c_int]),
    "drnh_grid_donor_info": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _I64P]),
    "drnh_grid_num_concat": (C.c_int64, [C.c_void_p, C.c_int, C.c_int]),
    "drnh_grid_rec_concat": (C.c_int, [C.c_void_p, C.c_int, _I64P]),
    "drnh_grid_donor_index": (C.c_int, [C.c_void_p, C.c_int, _I64P]),
    "drnh_grid_donor_weight": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "drnh_grid_donor_turn": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP]),
    "drnh_grid_set_base": (C.c_int, [C.c_void_p, C.c_int, _DP, _DP]),
    "drnh_grid_patch_base": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "drnh_grid_patch_transform": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "drnh_grid_levelset_spheres": (C.c_int, [C.c_void_p, C.c_int, _DP, C.c_int, C.c_int, C.c_double]),
    "drnh_grid_levelset_values": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "drnh_grid_ls_num_cells": (C.c_int64, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "drnh_grid_ls_cells": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _I64P, _DP, _I64P, _DP]),
    "drnh_grid_min_ch_length": (C.c_double, [C.c_void_p]),
    "drnh_dnd_write": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_void_p), _I64P]),
    "drnh_dnd_read": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_void_p), _I64P,
                                C.POINTER(C.c_int)]),
    "drnh_codestring": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int]),
    "drnh_codestring_compare": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_int32), C.c_char_p, C.c_int]),
    "drnh_grid_patch_codes": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int]),
    "drnh_grid_num_patch_groups": (C.c_int, [C.c_void_p]),
    "drnh_grid_patch_group": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p,
                                        C.c_int]),
    "drnh_vtk_write_multiblock": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "drnh_vtk_write_patch": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _DP, _DP, C.c_int,
                                       C.POINTER(C.c_char_p), C.POINTER(C.POINTER(C.c_float)), C.c_int,
                                       C.POINTER(C.c_char_p), C.POINTER(C.POINTER(C.c_float))]),
    "drnh_grid_to_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "drnh_grid_upload_exchange": (C.c_int, [C.c_void_p, C.c_void_p]),
    "drnh_grid_upload_levelset": (C.c_int, [C.c_void_p, C.c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdrnum.so is not built (%s); run python -m drnum_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _string_out(call, what, cap=256):
    """Run call(buf, cap) of a drnh_* string getter, growing the buffer on -2."""
    while True:
        buf = C.create_string_buffer(cap)
        rc = call(buf, cap)
        if rc == -2:
            cap *= 4
            continue
        if rc < 0:
            raise RuntimeError("%s failed (%d)" % (what, rc))
        return buf.value.decode()


def codestring(text: str) -> str:
    """CodeString(text) (codestring.cpp:33-103): the normalised code string."""
    return _string_out(lambda b, n: lib().drnh_codestring(text.encode(), b, n), "drnh_codestring")


def codestring_compare(a: str, b: str):
    """CodeString(a).compareCodeString(CodeString(b)) (codestring.cpp:106-174):
    (allsame, overruled, underruled, merged string)."""
    flags = (C.c_int32 * 3)()
    m = _string_out(lambda buf, n: lib().drnh_codestring_compare(a.encode(), b.encode(), flags, buf, n),
                    "drnh_codestring_compare")
    return bool(flags[0]), bool(flags[1]), bool(flags[2]), m


def vtk_write_multiblock(base: str, count: int, n_patches: int) -> str:
    """The .vtm index of PatchGrid::writeToVtk (patchgrid.cpp:673-695); returns
    the block directory (base with the count suffix)."""
    return _string_out(lambda b, n: lib().drnh_vtk_write_multiblock(base.encode(), count, n_patches, b, n),
                       "drnh_vtk_write_multiblock", cap=len(base) + 64)


def vtk_write_patch(path: str, n3, seek6, origin3, delta3, scalars: dict, vectors: dict):
    """One patch's core cells as a .vts block (CartesianPatch::createVtkDataSet,
    cartesianpatch.cpp:1661-1730): scalars name -> (N,), vectors name -> (3, N)
    in the patch layout."""
    n = np.ascontiguousarray(n3, dtype=np.int32)
    sk = np.ascontiguousarray(seek6, dtype=np.int32)
    o = np.ascontiguousarray(origin3, dtype=np.float64)
    d = np.ascontiguousarray(delta3, dtype=np.float64)
    fp = C.POINTER(C.c_float)
    sc = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in scalars.values()]
    vc = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in vectors.values()]
    sn = (C.c_char_p * max(len(sc), 1))(*[k.encode() for k in scalars])
    vn = (C.c_char_p * max(len(vc), 1))(*[k.encode() for k in vectors])
    sp = (fp * max(len(sc), 1))(*[a.ctypes.data_as(fp) for a in sc])
    vp = (fp * max(len(vc), 1))(*[a.ctypes.data_as(fp) for a in vc])
    rc = lib().drnh_vtk_write_patch(path.encode(), _ptr(n, C.c_int32), _ptr(sk, C.c_int32), _ptr(o, C.c_double),
                                    _ptr(d, C.c_double), len(sc), sn, sp, len(vc), vn, vp)
    if rc:
        raise RuntimeError("drnh_vtk_write_patch failed (%d)" % rc)


class HostGrid:
    """The drop-in's PatchGrid: patches, face functors, donor tables, level set."""

    def __init__(self, fp64_geometry: bool = True):
        self.L = lib()
        self.h = self.L.drnh_grid_create(1 if fp64_geometry else 0)
        self.fp64 = fp64_geometry

    def __del__(self):
        try:
            if self.h:
                self.L.drnh_grid_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            m = self.L.drnh_last_error(self.h)
            raise RuntimeError("%s failed: %s" % (what, m.decode() if m else ""))
        return rc

    @classmethod
    def from_workload(cls, w: configs.Workload, fp64_geometry: bool = True) -> "HostGrid":
        g = cls(fp64_geometry)
        for P in w.patches:
            seek = np.array(P.seek, dtype=np.int32)
            o = np.array(P.origin, dtype=np.float64)
            L = np.array(P.length, dtype=np.float64)
            pid = g._chk(g.L.drnh_grid_add_patch(g.h, P.ni, P.nj, P.nk, _ptr(seek, C.c_int32), _ptr(o, C.c_double),
                                                 _ptr(L, C.c_double)), "drnh_grid_add_patch")
            g.set_faces(pid, P.bc, P.bc_par)
            if not getattr(P, "aligned", True):
                g.set_base(pid, P.base_i, P.base_j)
        return g

    def set_base(self, p, base_i, base_j):
        bi = np.array(base_i, dtype=np.float64)
        bj = np.array(base_j, dtype=np.float64)
        self._chk(self.L.drnh_grid_set_base(self.h, p, _ptr(bi, C.c_double), _ptr(bj, C.c_double)), "set_base")

    def patch_base(self, p):
        out = np.zeros(6, np.float64)
        self._chk(self.L.drnh_grid_patch_base(self.h, p, _ptr(out, C.c_double)), "patch_base")
        return out[:3], out[3:]

    def patch_transform(self, p):
        """patch p's this->inertial transform as Patch::xyzoCell applies it: (A, b)"""
        out = np.zeros(12, np.float64)
        self._chk(self.L.drnh_grid_patch_transform(self.h, p, _ptr(out, C.c_double)), "patch_transform")
        return out[:9].reshape(3, 3), out[9:]

    def donor_turn(self, p, d):
        out = np.zeros(9, np.float64)
        self._chk(self.L.drnh_grid_donor_turn(self.h, p, d, _ptr(out, C.c_double)), "donor_turn")
        return out.reshape(3, 3)

    def read(self, path, scale=1.0):
        self._chk(self.L.drnh_grid_read(self.h, path.encode(), scale), "drnh_grid_read")

    def patch_codes(self, p) -> str:
        """The code string of patch p as read from the grid file."""
        return _string_out(lambda b, n: self.L.drnh_grid_patch_codes(self.h, p, b, n), "drnh_grid_patch_codes")

    def patch_groups(self):
        """PatchGroups of the grid (patchgroups.cpp:35-101), in creation order:
        dicts with type, patches (grid indices), codes, exceptions."""
        out = []
        for g in range(self.L.drnh_grid_num_patch_groups(self.h)):
            info = (C.c_int32 * 3)()
            codes = _string_out(lambda b, n: self.L.drnh_grid_patch_group(self.h, g, info, None, b, n),
                                "drnh_grid_patch_group")
            members = np.zeros(max(info[1], 1), np.int32)
            _string_out(lambda b, n: self.L.drnh_grid_patch_group(self.h, g, info, _ptr(members, C.c_int32), b, n),
                        "drnh_grid_patch_group")
            out.append(dict(type=int(info[0]), patches=members[:info[1]].tolist(), codes=codes,
                            exceptions=bool(info[2])))
        return out

    def num_patches(self):
        return self.L.drnh_grid_num_patches(self.h)

    def patch_info(self, p):
        oi = np.zeros(9, np.int32)
        od = np.zeros(9, np.float64)
        self._chk(self.L.drnh_grid_patch_info(self.h, p, _ptr(oi, C.c_int32), _ptr(od, C.c_double)), "patch_info")
        return dict(n=oi[:3].tolist(), seek=oi[3:].tolist(), origin=od[:3].tolist(), length=od[3:6].tolist(),
                    delta=od[6:].tolist())

    def set_faces(self, p, bc, par):
        b = np.ascontiguousarray(bc, dtype=np.int32)
        q = np.ascontiguousarray(par, dtype=np.float64).reshape(-1)
        self._chk(self.L.drnh_grid_set_faces(self.h, p, _ptr(b, C.c_int32), _ptr(q, C.c_double)), "set_faces")

    def compute_dependencies(self, period=(0.0, 0.0, 0.0), threads: int = 8):
        per = np.array(period, dtype=np.float64)
        self._chk(self.L.drnh_grid_compute_dependencies(self.h, _ptr(per, C.c_double), threads),
                  "drnh_grid_compute_dependencies")

    def tables(self, p):
        L, h = self.L, self.h
        nrec = L.drnh_grid_num_receivers(h, p)
        recv = np.zeros(max(nrec, 1), np.int64)
        L.drnh_grid_receivers(h, p, _ptr(recv, C.c_int64))
        nd = L.drnh_grid_num_donors(h, p)
        donors = []
        info = np.zeros(5, np.int64)
        for d in range(nd):
            L.drnh_grid_donor_info(h, p, d, _ptr(info, C.c_int64))
            donors.append(tuple(int(x) for x in info))
        nrc = L.drnh_grid_num_concat(h, p, 0)
        nwi = L.drnh_grid_num_concat(h, p, 1)
        rc = np.zeros(max(nrc, 1), np.int64)
        di = np.zeros(max(nwi, 1), np.int64)
        dw = np.zeros(max(nwi, 1), np.float64)
        L.drnh_grid_rec_concat(h, p, _ptr(rc, C.c_int64))
        L.drnh_grid_donor_index(h, p, _ptr(di, C.c_int64))
        L.drnh_grid_donor_weight(h, p, _ptr(dw, C.c_double))
        turn = np.array([self.donor_turn(p, d).ravel() for d in range(nd)]).reshape(nd, 9)
        return dict(receivers=recv[:nrec], donors=donors, rec_concat=rc[:nrc], donor_index=di[:nwi],
                    donor_weight=dw[:nwi], no_donor=int(L.drnh_grid_num_no_donor(h, p)), turn=turn)

    def levelset_spheres(self, spheres, n_inner=3, n_outer=0, min_inner_rel=0.1):
        s = np.ascontiguousarray(spheres, dtype=np.float64).reshape(-1)
        self._chk(self.L.drnh_grid_levelset_spheres(self.h, s.size // 4, _ptr(s, C.c_double), n_inner, n_outer,
                                                    min_inner_rel), "drnh_grid_levelset_spheres")

    def levelset_values(self, p, n):
        out = np.zeros(n, np.float64)
        self._chk(self.L.drnh_grid_levelset_values(self.h, p, _ptr(out, C.c_double)), "levelset_values")
        return out

    def ls_cells(self, p, layer, outer=False):
        n = self.L.drnh_grid_ls_num_cells(self.h, p, layer, int(outer))
        cells = np.zeros(max(n, 1), np.int64)
        g = np.zeros((max(n, 1), 4), np.float64)
        donors = np.zeros((max(n, 1), 8), np.int64)
        w = np.zeros((max(n, 1), 8), np.float64)
        self._chk(self.L.drnh_grid_ls_cells(self.h, p, layer, int(outer), _ptr(cells, C.c_int64), _ptr(g, C.c_double),
                                            _ptr(donors, C.c_int64), _ptr(w, C.c_double)), "ls_cells")
        return dict(cells=cells[:n], g=g[:n], donors=donors[:n], weights=w[:n])

    def min_ch_length(self):
        return self.L.drnh_grid_min_ch_length(self.h)

    # ---- hand-over to a device context ------------------------------------
    def to_device(self, ctx: hip.Context):
        self._chk(self.L.drnh_grid_to_device(self.h, ctx.h), "drnh_grid_to_device")

    def upload_exchange(self, ctx: hip.Context):
        self._chk(self.L.drnh_grid_upload_exchange(self.h, ctx.h), "drnh_grid_upload_exchange")

    def upload_levelset(self, ctx: hip.Context):
        self._chk(self.L.drnh_grid_upload_levelset(self.h, ctx.h), "drnh_grid_upload_levelset")


# ---- restart files ------------------------------------------------------------

def dnd_write(path, fields, time, precision=None):
    """Write one field of every patch (list of (5, N) arrays) as a DrNUM .dnd file."""
    prec = precision or (4 if np.asarray(fields[0]).dtype == np.float32 else 8)
    dt = np.float32 if prec == 4 else np.float64
    arrs = [np.ascontiguousarray(f, dtype=dt).reshape(-1) for f in fields]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    counts = np.array([a.size for a in arrs], dtype=np.int64)
    rc = lib().drnh_dnd_write(path.encode(), prec, float(time), len(arrs), ptrs, _ptr(counts, C.c_int64))
    if rc != 0:
        raise RuntimeError("drnh_dnd_write failed (%d)" % rc)


def dnd_read(path, shapes, dtype=np.float64):
    """Read a .dnd file into arrays of the given (5, N) shapes; returns (fields, time, file_precision)."""
    dt = np.dtype(dtype)
    prec = 4 if dt == np.float32 else 8
    arrs = [np.zeros(s, dtype=dt) for s in shapes]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    counts = np.array([a.size for a in arrs], dtype=np.int64)
    t = C.c_double()
    fp = C.c_int()
    rc = lib().drnh_dnd_read(path.encode(), prec, C.byref(t), len(arrs), ptrs, _ptr(counts, C.c_int64), C.byref(fp))
    if rc != 0:
        raise RuntimeError("drnh_dnd_read failed (%d)" % rc)
    return arrs, t.value, fp.value
