"""Perfect-gas model of the reference (perfectgas.h:31-98), vectorised in numpy.

The constants are the reference's: R = 287.2, gamma = 1.4, cp = 1004.5,
cv = 717.5 (note R/cv = 0.40028 differs from gamma - 1; the reference uses
gamma only in the sound speed).  `prim_to_cons` reproduces the reference's
mixed precision: in the float build the energy term is evaluated in double
because of the literal 0.5 (perfectgas.h:63).
"""
import numpy as np

R = 287.2
GAMMA = 1.4
CP = 1004.5
CV = 717.5


def prim_to_cons(p, T, u, v, w, dtype=np.float64) -> np.ndarray:
    """perfectgas.h:56-64 PerfectGas::primitiveToConservative(p, T, u, v, w)."""
    dt = np.dtype(dtype)
    f = dt.type
    p, T, u, v, w = (np.asarray(a, dtype=dt) for a in (p, T, u, v, w))
    r = p / (f(R) * T)
    ke = u * u + v * v + w * w
    rE = (r.astype(np.float64) * ((f(CV) * T).astype(np.float64) + 0.5 * ke.astype(np.float64))).astype(dt)
    return np.stack([r, r * u, r * v, r * w, rE])


def cons_to_prim(var: np.ndarray):
    """perfectgas.h:83-92 PerfectGas::conservativeToPrimitive -> p, T, u, v, w (float64)."""
    var = np.asarray(var, dtype=np.float64)
    ir = 1.0 / var[0]
    u, v, w = var[1] * ir, var[2] * ir, var[3] * ir
    T = (var[4] * ir - 0.5 * (u * u + v * v + w * w)) / CV
    p = var[0] * R * T
    return p, T, u, v, w


def sound_speed(T):
    return np.sqrt(GAMMA * R * np.asarray(T, dtype=np.float64))
