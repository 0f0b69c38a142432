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


# ---------------------------------------------------------------- hydro
# The hydro build (no field): Euler equations in the same 9-slot layout
# with the field slots zero.  Written from E. F. Toro, "Riemann Solvers and
# Numerical Methods for Fluid Dynamics", 3rd ed. (2009).

def sound_speed(w, gam):
    return math.sqrt(gam * w[4] / w[0])


def hll_hydro(wl, wr, gam):
    """HLL with the simple speed bounds min(u) - max(a), max(u) + max(a)."""
    a = max(sound_speed(wl, gam), sound_speed(wr, gam))
    sl, sr = min(wl[1], wr[1]) - a, max(wl[1], wr[1]) + a
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


def pvrs_speeds(wl, wr, gam):
    """Pressure-based wave speed estimates (Toro sec. 10.5.2, eqs. 10.59-10.61):
    p* from the primitive-variable Riemann solver, then S_K = u_K -/+ a_K q_K."""
    al, ar = sound_speed(wl, gam), sound_speed(wr, gam)
    rho_bar, a_bar = 0.5 * (wl[0] + wr[0]), 0.5 * (al + ar)
    pstar = max(0.0, 0.5 * (wl[4] + wr[4]) - 0.5 * (wr[1] - wl[1]) * rho_bar * a_bar)

    def q(p):
        return 1.0 if pstar <= p else math.sqrt(1.0 + (gam + 1.0) / (2.0 * gam) * (pstar / p - 1.0))

    return wl[1] - al * q(wl[4]), wr[1] + ar * q(wr[4])


def hllc_toro(wl, wr, gam):
    """HLLC (Toro sec. 10.4): S* from eq. (10.37), star states U*_K from
    eq. (10.39), F*_K = F_K + S_K (U*_K - U_K) (10.38); the passive scalar
    is advected like a tangential velocity."""
    sl, sr = pvrs_speeds(wl, wr, gam)
    rl, ul, pl = wl[0], wl[1], wl[4]
    rr, ur, pr = wr[0], wr[1], wr[4]
    ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))
    Ul, Ur = prim_to_cons(wl, gam), prim_to_cons(wr, gam)
    Fl, Fr = phys_flux(wl, gam), phys_flux(wr, gam)

    def ustar(w, U, s):
        r, u, p = w[0], w[1], w[4]
        c = r * (s - u) / (s - ss)
        out = np.zeros(9)
        out[0] = c
        out[1] = c * ss
        out[2] = c * w[2]
        out[3] = c * w[3]
        out[4] = c * (U[4] / r + (ss - u) * (ss + p / (r * (s - u))))
        out[5] = c * w[5] / r
        return out

    if sl >= 0:
        f = Fl
    elif ss >= 0:
        f = Fl + sl * (ustar(wl, Ul, sl) - Ul)
    elif sr > 0:
        f = Fr + sr * (ustar(wr, Ur, sr) - Ur)
    else:
        f = Fr
    f = np.array(f, dtype=float)
    f[6:9] = 0.0
    return f


def random_hydro_states(rng, n, vmax=1.0):
    """Random primitive pairs without field."""
    wl = np.zeros((9, n))
    wr = np.zeros((9, n))
    for w in (wl, wr):
        w[0] = rng.uniform(0.1, 2.0, n)
        w[1:4] = rng.uniform(-vmax, vmax, (3, n))
        w[4] = rng.uniform(0.05, 2.0, n)
        w[5] = rng.uniform(0.0, 1.0, n)
    return wl, wr, np.zeros(n)


def exact_riemann_euler(left, right, gam, xi):
    """Exact solution of the Riemann problem of the Euler equations (Toro
    ch. 4): left, right = (rho, u, p); returns rho, u, p sampled at the
    similarity coordinates xi = (x - x0)/t.  p* by Newton iteration on
    f_L(p) + f_R(p) + u_R - u_L = 0 (eqs. 4.5-4.7)."""
    rl, ul, pl = left
    rr, ur, pr = right
    al, ar = math.sqrt(gam * pl / rl), math.sqrt(gam * pr / rr)

    def fk(p, r, pk, a):
        if p > pk:  # shock
            A, B = 2.0 / ((gam + 1.0) * r), (gam - 1.0) / (gam + 1.0) * pk
            f = (p - pk) * math.sqrt(A / (p + B))
            d = math.sqrt(A / (B + p)) * (1.0 - 0.5 * (p - pk) / (B + p))
        else:  # rarefaction
            f = 2.0 * a / (gam - 1.0) * ((p / pk) ** ((gam - 1.0) / (2.0 * gam)) - 1.0)
            d = (p / pk) ** (-(gam + 1.0) / (2.0 * gam)) / (r * a)
        return f, d

    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        f1, d1 = fk(p, rl, pl, al)
        f2, d2 = fk(p, rr, pr, ar)
        pn = max(1e-12, p - (f1 + f2 + ur - ul) / (d1 + d2))
        if abs(pn - p) < 1e-14 * p:
            p = pn
            break
        p = pn
    ps = p
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, ar)[0] - fk(ps, rl, pl, al)[0])
    g1, g2 = (gam - 1.0) / (gam + 1.0), (gam - 1.0) / (2.0 * gam)
    out = np.zeros((3, len(xi)))
    for n, s in enumerate(xi):
        if s <= us:  # left of the contact
            if ps > pl:
                sh = ul - al * math.sqrt((gam + 1.0) / (2.0 * gam) * ps / pl + g2)
                out[:, n] = (rl, ul, pl) if s <= sh else (rl * (ps / pl + g1) / (g1 * ps / pl + 1.0), us, ps)
            else:
                ast = al * (ps / pl) ** g2
                if s <= ul - al:
                    out[:, n] = (rl, ul, pl)
                elif s >= us - ast:
                    out[:, n] = (rl * (ps / pl) ** (1.0 / gam), us, ps)
                else:
                    c = 2.0 / (gam + 1.0) + g1 / al * (ul - s)
                    out[:, n] = (rl * c ** (2.0 / (gam - 1.0)),
                                 2.0 / (gam + 1.0) * (al + (gam - 1.0) / 2.0 * ul + s),
                                 pl * c ** (2.0 * gam / (gam - 1.0)))
        else:
            if ps > pr:
                sh = ur + ar * math.sqrt((gam + 1.0) / (2.0 * gam) * ps / pr + g2)
                out[:, n] = (rr, ur, pr) if s >= sh else (rr * (ps / pr + g1) / (g1 * ps / pr + 1.0), us, ps)
            else:
                ast = ar * (ps / pr) ** g2
                if s >= ur + ar:
                    out[:, n] = (rr, ur, pr)
                elif s <= us + ast:
                    out[:, n] = (rr * (ps / pr) ** (1.0 / gam), us, ps)
                else:
                    c = 2.0 / (gam + 1.0) - g1 / ar * (ur - s)
                    out[:, n] = (rr * c ** (2.0 / (gam - 1.0)),
                                 2.0 / (gam + 1.0) * (-ar + (gam - 1.0) / 2.0 * ur + s),
                                 pr * c ** (2.0 * gam / (gam - 1.0)))
    return out
