"""Input files of the reference's formats, written from a workload description."""
import numpy as np

from drnum_amd import configs


def grid_file_text(w):
    """A grid file as the reference's Python writer lays it out (drnum.py:199-244)."""
    codes = {configs.BC_FARFIELD: "far", configs.BC_SLIP: "slip", configs.BC_TRANSMISSIVE: "trans",
             configs.BC_JET: "jet", configs.BC_NONE: "none"}
    out = []
    for n, P in enumerate(w.patches):
        txt = "1001 // index=%d name='p%d'\n{\n" % (n, n)
        txt += "  %r %r %r\n" % tuple(P.origin)
        txt += "  %r %r %r\n" % tuple(float(b) for b in P.base_i)
        txt += "  %r %r %r\n" % tuple(float(b) for b in P.base_j)
        txt += "  1\n"
        txt += "  %d %d %d\n" % (P.ni, P.nj, P.nk)
        txt += "  " + " ".join(str(s) for s in P.seek) + "\n"
        txt += "  %r %r %r\n" % tuple(P.length)
        txt += "  fx fy fz\n"
        txt += "  " + " ".join("0" if P.seek[f] else codes[P.bc[f]] for f in range(6)) + "\n"
        txt += "  0\n}\n"
        out.append(txt)
    out.append("0\n")
    return "".join(out)


def case_file_bytes(w, dtype, levelset=None):
    """The case file of drnum_run: per patch int32 bc[6], float64 par[6][16] and
    the initial field 0 (5*N values of `dtype`); optionally the level-set spheres."""
    parts = []
    for p, P in enumerate(w.patches):
        parts.append(np.asarray(P.bc, dtype=np.int32).tobytes())
        parts.append(np.ascontiguousarray(P.bc_par, dtype=np.float64).tobytes())
        parts.append(np.ascontiguousarray(w.initial_state(p, dtype)).tobytes())
    if levelset is not None:
        spheres, n_inner, min_rel = levelset
        s = np.ascontiguousarray(spheres, dtype=np.float64)
        parts.append(np.array([len(s), n_inner], dtype=np.int32).tobytes())
        parts.append(np.array([min_rel], dtype=np.float64).tobytes())
        parts.append(s.tobytes())
    return b"".join(parts)
