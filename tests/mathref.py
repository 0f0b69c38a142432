"""Independent textbook formulas used to pin the oracle.

Written from the published methods, not from the reference's code:
  * ideal-MHD physical flux and fast magnetosonic speed (any MHD text, e.g.
    Toro, "Riemann Solvers and Numerical Methods", ch. 1 and 10);
  * HLL (Harten, Lax & van Leer 1983) with Davis-type speed bounds;
  * HLLD (Miyoshi & Kusano 2005, J. Comput. Phys. 208, 315), eqs. (38), (41),
    (43)-(48), (51), (59)-(63) in the paper's own form: S_M and p_T* from the
    two-sided jump relations rather than from the HLL average state the
    reference uses (src/mhd/fluxes_c.c:668-697), so agreement is a check of
    the algebra, not a retyping of it.
Vectors are primitive (rho, vx, vy, vz, p, psc, bx, by, bz) and conserved
(rho, mx, my, mz, E, rho*psc... see below) in the reference's slot order.
The passive scalar is carried as the conserved density u[5]=w[5] exactly as
the reference does (src/mhd/eos_c.h:87).
"""
import math

import numpy as np


def prim_to_cons(w, gam):
    rho, vx, vy, vz, p, s, bx, by, bz = w
    e = p / (gam - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    return np.array([rho, rho * vx, rho * vy, rho * vz, e, s, bx, by, bz])


def phys_flux(w, gam):
    """x-flux of ideal MHD with total pressure p + B^2/2."""
    rho, vx, vy, vz, p, s, bx, by, bz = w
    u = prim_to_cons(w, gam)
    pt = p + 0.5 * (bx * bx + by * by + bz * bz)
    vb = vx * bx + vy * by + vz * bz
    return np.array([
        rho * vx,
        rho * vx * vx + pt - bx * bx,
        rho * vx * vy - bx * by,
        rho * vx * vz - bx * bz,
        (u[4] + pt) * vx - bx * vb,
        s * vx,
        0.0,
        by * vx - bx * vy,
        bz * vx - bx * vz,
    ])


def fast_speed(w, gam, bn=None):
    """Fast magnetosonic speed along the axis whose field component is bn."""
    rho, p = w[0], w[4]
    b2 = w[6] ** 2 + w[7] ** 2 + w[8] ** 2
    bn = w[6] if bn is None else bn
    a2 = gam * p / rho
    va2 = b2 / rho
    return math.sqrt(0.5 * (a2 + va2 + math.sqrt(max((a2 + va2) ** 2 - 4 * a2 * bn * bn / rho, 0.0))))


def speeds(wl, wr, gam):
    cf = max(fast_speed(wl, gam), fast_speed(wr, gam))
    return min(wl[1], wr[1]) - cf, max(wl[1], wr[1]) + cf


def hll(wl, wr, gam):
    sl, sr = speeds(wl, wr, gam)
    ul, ur = prim_to_cons(wl, gam), prim_to_cons(wr, gam)
    fl, fr = phys_flux(wl, gam), phys_flux(wr, gam)
    if sl >= 0:
        f = fl
    elif sr <= 0:
        f = fr
    else:
        f = (sr * fl - sl * fr + sl * sr * (ur - ul)) / (sr - sl)
    f = f.copy()
    f[6] = 0.0
    return f


def hlld(wl, wr, gam, return_states=False):
    """Miyoshi & Kusano (2005) HLLD flux; wl[6] and wr[6] must be equal."""
    bx = wl[6]
    sl, sr = speeds(wl, wr, gam)
    ul, ur = prim_to_cons(wl, gam), prim_to_cons(wr, gam)
    fl, fr = phys_flux(wl, gam), phys_flux(wr, gam)
    rl, rr = wl[0], wr[0]
    ul_, ur_ = wl[1], wr[1]
    ptl = wl[4] + 0.5 * (bx * bx + wl[7] ** 2 + wl[8] ** 2)
    ptr = wr[4] + 0.5 * (bx * bx + wr[7] ** 2 + wr[8] ** 2)
    dl = (sl - ul_) * rl
    dr = (sr - ur_) * rr
    sm = (dr * ur_ - dl * ul_ - ptr + ptl) / (dr - dl)                      # (38)
    pts = (dr * ptl - dl * ptr + rl * rr * (sr - ur_) * (sl - ul_) * (ur_ - ul_)) / (dr - dl)  # (41)
    sgn = float(np.sign(bx))

    def star(w, u, s):
        rho, vn = w[0], w[1]
        rs = rho * (s - vn) / (s - sm)                                     # (43)
        den = rho * (s - vn) * (s - sm) - bx * bx
        vy = w[2] - bx * w[7] * (sm - vn) / den                            # (44)
        vz = w[3] - bx * w[8] * (sm - vn) / den
        by = w[7] * (rho * (s - vn) ** 2 - bx * bx) / den                  # (45)
        bz = w[8] * (rho * (s - vn) ** 2 - bx * bx) / den
        pt = w[4] + 0.5 * (bx * bx + w[7] ** 2 + w[8] ** 2)
        vb = w[1] * bx + w[2] * w[7] + w[3] * w[8]
        vbs = sm * bx + vy * by + vz * bz
        e = ((s - vn) * u[4] - pt * vn + pts * sm + bx * (vb - vbs)) / (s - sm)   # (48)
        return dict(rho=rs, vy=vy, vz=vz, by=by, bz=bz, e=e, psc=u[5] * (s - vn) / (s - sm))

    def vec(st, vy, vz, by, bz, e):
        r = st["rho"]
        return np.array([r, r * sm, r * vy, r * vz, e, st["psc"], bx, by, bz])

    L, R = star(wl, ul, sl), star(wr, ur, sr)
    usl = vec(L, L["vy"], L["vz"], L["by"], L["bz"], L["e"])
    usr = vec(R, R["vy"], R["vz"], R["by"], R["bz"], R["e"])
    sql, sqr_ = math.sqrt(L["rho"]), math.sqrt(R["rho"])
    ssl = sm - abs(bx) / sql                                                # (51)
    ssr = sm + abs(bx) / sqr_
    den = sql + sqr_
    vyss = (sql * L["vy"] + sqr_ * R["vy"] + (R["by"] - L["by"]) * sgn) / den   # (59)
    vzss = (sql * L["vz"] + sqr_ * R["vz"] + (R["bz"] - L["bz"]) * sgn) / den
    byss = (sql * R["by"] + sqr_ * L["by"] + sql * sqr_ * (R["vy"] - L["vy"]) * sgn) / den  # (61)
    bzss = (sql * R["bz"] + sqr_ * L["bz"] + sql * sqr_ * (R["vz"] - L["vz"]) * sgn) / den
    vbss = sm * bx + vyss * byss + vzss * bzss
    vbl = sm * bx + L["vy"] * L["by"] + L["vz"] * L["bz"]
    vbr = sm * bx + R["vy"] * R["by"] + R["vz"] * R["bz"]
    essl = L["e"] - sql * (vbl - vbss) * sgn                               # (63)
    essr = R["e"] + sqr_ * (vbr - vbss) * sgn
    ussl = vec(L, vyss, vzss, byss, bzss, essl)
    ussr = vec(R, vyss, vzss, byss, bzss, essr)
    fsl = fl + sl * (usl - ul)
    fsr = fr + sr * (usr - ur)
    fssl = fsl + ssl * (ussl - usl)
    fssr = fsr + ssr * (ussr - usr)
    if sl > 0:
        f = fl
    elif ssl > 0:
        f = fsl
    elif sm > 0:
        f = fssl
    elif ssr > 0:
        f = fssr
    elif sr > 0:
        f = fsr
    else:
        f = fr
    f = f.copy()
    f[6] = 0.0
    if return_states:
        return f, dict(sl=sl, sr=sr, sm=sm, ssl=ssl, ssr=ssr, usl=usl, usr=usr,
                       ussl=ussl, ussr=ussr, ul=ul, ur=ur, fl=fl, fr=fr)
    return f


def random_states(rng, n, bx_min=0.1, vmax=1.0):
    """Random primitive pairs away from HLLD's degenerate cases."""
    wl = np.zeros((9, n))
    wr = np.zeros((9, n))
    for w in (wl, wr):
        w[0] = rng.uniform(0.1, 2.0, n)
        w[1:4] = rng.uniform(-vmax, vmax, (3, n))
        w[4] = rng.uniform(0.05, 2.0, n)
        w[5] = rng.uniform(0.0, 1.0, n)
        w[7:9] = rng.uniform(-1.5, 1.5, (2, n))
    bx = rng.uniform(bx_min, 1.5, n) * rng.choice([-1.0, 1.0], n)
    wl[6] = bx
    wr[6] = bx
    return wl, wr, bx
