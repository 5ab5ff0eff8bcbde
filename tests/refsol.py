"""Independent reference solutions used to pin the oracle.

Nothing here restates the oracle's code.  Each routine follows a published
method or a closed form:

  exact_riemann   the exact solution of the 1-D Riemann problem for the Euler
                  equations of a perfect gas (Toro, "Riemann Solvers and
                  Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4:
                  pressure function, Newton iteration for p*, wave sampling).
  ausm_plus       AUSM+ as published (Liou, J. Comput. Phys. 129 (1996) 364,
                  eqs. 19-21 and 25-26): fourth-order split Mach polynomial
                  with beta = 1/8, fifth-order split pressure with
                  alpha = 3/16, interface flux
                  a_1/2 * (m+ rho phi_L + m- rho phi_R) + p_1/2 n, with the
                  arithmetic-mean interface sound speed that DrNUM's
                  AusmPlus uses (ausmplus.h:55).
  euler_flux      the physical Euler flux F(U) . n.

The gas is DrNUM's PerfectGas (perfectgas.h:31-34): p = rho R T with
T = (E/rho - |u|^2/2)/cv, so the PDE is the Euler system with
gamma_eff = 1 + R/cv, while the sound speed in the flux splitting uses
gamma = 1.4.
"""
from __future__ import annotations

import numpy as np

R = 287.2
GAMMA = 1.4
CV = 717.5
CP = 1004.5
GAMMA_EFF = 1.0 + R / CV


# ---------------------------------------------------------------------------
# conserved <-> primitive with the PerfectGas law
# ---------------------------------------------------------------------------

def primitives(U):
    """U = (rho, rho u, rho v, rho w, rho E) -> rho, (u, v, w), p, T, H."""
    U = np.asarray(U, dtype=np.float64)
    rho = U[0]
    vel = U[1:4] / rho
    T = (U[4] / rho - 0.5 * np.sum(vel * vel, axis=0)) / CV
    p = rho * R * T
    H = (U[4] + p) / rho
    return rho, vel, p, T, H


def euler_flux(U, n, area=1.0):
    rho, vel, p, T, H = primitives(U)
    un = vel[n]
    F = np.empty(5)
    F[0] = rho * un
    F[1:4] = rho * un * vel
    F[1 + n] += p
    F[4] = rho * un * H
    return area * F


# ---------------------------------------------------------------------------
# AUSM+ (Liou 1996)
# ---------------------------------------------------------------------------

def _m2(M, sgn):
    return sgn * 0.25 * (M + sgn) ** 2


def _m4(M, sgn, beta=1.0 / 8.0):
    if abs(M) >= 1.0:
        return 0.5 * (M + sgn * abs(M))
    return _m2(M, sgn) * (1.0 - sgn * 16.0 * beta * _m2(M, -sgn))


def _p5(M, sgn, alpha=3.0 / 16.0):
    if abs(M) >= 1.0:
        return 0.5 * (M + sgn * abs(M)) / M
    return _m2(M, sgn) * ((sgn * 2.0 - M) - sgn * 16.0 * alpha * M * _m2(M, -sgn))


def ausm_plus(UL, UR, n, area=1.0):
    rl, vl, pl, Tl, Hl = primitives(UL)
    rr, vr, pr, Tr, Hr = primitives(UR)
    al = np.sqrt(GAMMA * R * Tl)
    ar = np.sqrt(GAMMA * R * Tr)
    a = 0.5 * (al + ar)
    ML = vl[n] / a
    MR = vr[n] / a
    m = _m4(ML, +1) + _m4(MR, -1)
    p_half = _p5(ML, +1) * pl + _p5(MR, -1) * pr
    mp = max(m, 0.0)
    mm = min(m, 0.0)
    phiL = np.array([1.0, vl[0], vl[1], vl[2], Hl]) * rl
    phiR = np.array([1.0, vr[0], vr[1], vr[2], Hr]) * rr
    F = a * (mp * phiL + mm * phiR)
    F[1 + n] += p_half
    return area * F


# ---------------------------------------------------------------------------
# exact Riemann solver (Toro ch. 4)
# ---------------------------------------------------------------------------

def _f_k(p, rk, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + B))
    # rarefaction
    e = (g - 1.0) / (2.0 * g)
    f = 2.0 * ck / (g - 1.0) * ((p / pk) ** e - 1.0)
    df = 1.0 / (rk * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _f_k(p, rl, pl, cl, g)
        fr, dfr = _f_k(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        p_new = max(1e-12, p - dp)
        if abs(p_new - p) <= 1e-15 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _f_k(p, rl, pl, cl, g)
    fr, _ = _f_k(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def exact_riemann(rl, ul, pl, rr, ur, pr, g, s):
    """Density, velocity, pressure at similarity coordinates s = (x - x0)/t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    out = np.zeros((3, len(s)))
    for n, S in enumerate(np.asarray(s, dtype=np.float64)):
        if S <= us:  # left of the contact
            if ps > pl:  # left shock
                SL = ul - cl * np.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                if S <= SL:
                    st = (rl, ul, pl)
                else:
                    st = (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:  # left rarefaction
                cls = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                SH, ST = ul - cl, us - cls
                if S <= SH:
                    st = (rl, ul, pl)
                elif S >= ST:
                    st = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - S))
                    u = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + S)
                    st = (rl * (c / cl) ** (2.0 / (g - 1.0)), u, pl * (c / cl) ** (2.0 * g / (g - 1.0)))
        else:  # right of the contact
            if ps > pr:  # right shock
                SR = ur + cr * np.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                if S >= SR:
                    st = (rr, ur, pr)
                else:
                    st = (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:  # right rarefaction
                crs = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                SH, ST = ur + cr, us + crs
                if S >= SH:
                    st = (rr, ur, pr)
                elif S <= ST:
                    st = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - S))
                    u = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + S)
                    st = (rr * (c / cr) ** (2.0 / (g - 1.0)), u, pr * (c / cr) ** (2.0 * g / (g - 1.0)))
        out[:, n] = st
    return out


# ---------------------------------------------------------------------------
# patch index helpers
# ---------------------------------------------------------------------------

def ijk(P, cell):
    """Local (i, j, k) of flat cell index cell = i*NJ*NK + j*NK + k."""
    cell = np.asarray(cell, dtype=np.int64)
    njk = P.nj * P.nk
    i = cell // njk
    r = cell - i * njk
    j = r // P.nk
    return i, j, r - j * P.nk


def flat(P, i, j, k):
    return (np.asarray(i, dtype=np.int64) * P.nj + j) * P.nk + k


def seek_cells(P):
    """All seek (receiving) cells of a patch in ascending flat order."""
    i, j, k = np.meshgrid(np.arange(P.ni), np.arange(P.nj), np.arange(P.nk), indexing="ij")
    s = P.seek
    m = ((i < s[0]) | (i >= P.ni - s[1]) | (j < s[2]) | (j >= P.nj - s[3]) | (k < s[4]) | (k >= P.nk - s[5]))
    return np.nonzero(m.ravel())[0].astype(np.int64)


def core_mask(P):
    i, j, k = np.meshgrid(np.arange(P.ni), np.arange(P.nj), np.arange(P.nk), indexing="ij")
    s = P.seek
    m = ((i >= s[0]) & (i < P.ni - s[1]) & (j >= s[2]) & (j < P.nj - s[3]) & (k >= s[4]) & (k < P.nk - s[5]))
    return m.ravel()
