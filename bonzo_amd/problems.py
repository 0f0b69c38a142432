"""Initial conditions for the configurations the MHD step is run on.

Each generator returns ``(prim, bfld)`` over the full grid of a
``GridShape`` (active cells set, ghosts left for the boundary fill), with the
reference's layout: ``prim`` is [9][Nz][Ny][Nx] with rows rho, vx, vy, vz, p,
passive scalar, Bx, By, Bz (cell-centred B is filled later by ``interp_bc``),
``bfld`` is [3][Nz][Ny][Nx] indexed by the lower face
(src/coordinates/grid.pyx:28-52, src/integration/ct.pyx:110-111).

Face-centred fields are built as discrete curls of an edge-centred vector
potential so that the discrete divergence is zero to round-off, as the
reference's Orszag-Tang setup does (src/problem/otvortex.pyx:38-66).
Random fields use numpy's PCG64 with an explicit seed.
"""
import math
from dataclasses import dataclass

import numpy as np

RHO, VX, VY, VZ, PR, PSC, BXC, BYC, BZC = range(9)


@dataclass(frozen=True)
class GridShape:
    """Active size, ghost count and box of a uniform Cartesian grid."""
    nx: int
    ny: int = 1
    nz: int = 1
    ng: int = 3
    lmin: tuple = (0.0, 0.0, 0.0)
    lmax: tuple = (1.0, 1.0, 1.0)

    @property
    def dims(self):
        return 3 if self.nz > 1 else (2 if self.ny > 1 else 1)

    @property
    def ntot(self):
        d = self.dims
        return (self.nx + 2 * self.ng,
                self.ny + 2 * self.ng if d >= 2 else 1,
                self.nz + 2 * self.ng if d >= 3 else 1)

    @property
    def nact(self):
        return (self.nx, self.ny, self.nz)

    def active(self):
        """Slices (k, j, i) of the active cells."""
        g = self.ng
        d = self.dims
        si = slice(g, g + self.nx)
        sj = slice(g, g + self.ny) if d >= 2 else slice(0, 1)
        sk = slice(g, g + self.nz) if d >= 3 else slice(0, 1)
        return sk, sj, si

    def dx(self, axis):
        return (self.lmax[axis] - self.lmin[axis]) / self.nact[axis]

    def centres(self, axis):
        """Cell-centre coordinates of the active cells along an axis.

        Uniform faces start at 0 regardless of lmin (src/coordinates/
        coord.pyx:205-209), so centres are (i+1/2)*dx."""
        n = self.nact[axis]
        return (np.arange(n) + 0.5) * self.dx(axis)

    def empty(self):
        nt = self.ntot
        prim = np.zeros((9, nt[2], nt[1], nt[0]))
        bfld = np.zeros((3, nt[2], nt[1], nt[0]))
        return prim, bfld


def _put(arr, shape, values):
    sk, sj, si = shape.active()
    arr[sk, sj, si] = values


def orszag_tang(n, ng=3, b_eps=1e-20):
    """2D Orszag-Tang vortex on [0,1]^2 in Heaviside-Lorentz-free units:
    rho=25/(36 pi), p=5/(12 pi), v=(-sin 2pi y, sin 2pi x), B from
    A_z = B0 (cos 4pi x/(4pi) + cos 2pi y/(2pi)) at cell corners with
    B0 = 1/sqrt(4 pi) (src/problem/otvortex.pyx:38-66 scaled by 1/(4 pi)).
    ``b_eps`` is the reference's 1e-20 offset that keeps face B non-zero
    (otvortex.pyx:65-66)."""
    s = GridShape(n, n, 1, ng)
    prim, bfld = s.empty()
    x = s.centres(0)
    y = s.centres(1)
    X, Y = np.meshgrid(x, y, indexing="xy")
    _put(prim[RHO], s, 25.0 / (36.0 * math.pi))
    _put(prim[PR], s, 5.0 / (12.0 * math.pi))
    _put(prim[VX], s, -np.sin(2 * math.pi * Y))
    _put(prim[VY], s, np.sin(2 * math.pi * X))
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    dx, dy = s.dx(0), s.dx(1)
    xe = np.arange(n + 1) * dx          # corner coordinates
    ye = np.arange(n + 1) * dy
    XE, YE = np.meshgrid(xe, ye, indexing="xy")
    az = b0 * (np.cos(4 * math.pi * XE) / (4 * math.pi) + np.cos(2 * math.pi * YE) / (2 * math.pi))
    bx = (az[1:, :-1] - az[:-1, :-1]) / dy + b_eps   # face (i-1/2, j)
    by = -(az[:-1, 1:] - az[:-1, :-1]) / dx + b_eps  # face (i, j-1/2)
    _put(bfld[0], s, bx)
    _put(bfld[1], s, by)
    return s, prim, bfld


def _band_modes(rng, slope, kmin, kmax, ncomp):
    """Seeded complex Gaussian coefficients c_k (unit variance of the real
    and imaginary parts) times |k|**slope on the shell band kmin <= |k| <=
    kmax of integer wavenumbers, zero elsewhere.  Returns (coef, kvec):
    coef has shape (ncomp, 2K+1, 2K+1, 2K+1) over (kz, ky, kx) in [-K, K]
    with K = floor(kmax), kvec the three wavenumber grids of that shape."""
    K = int(math.floor(kmax))
    k1 = np.arange(-K, K + 1, dtype=np.float64)
    KZ, KY, KX = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[band] = kk[band] ** slope
    shp = (ncomp,) + kk.shape
    coef = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * amp
    return coef, (KX, KY, KZ)


def _synth(coef, n, stagger):
    """Real field f(x) = Re sum_k coef_k exp(2 pi i k.x) of one coefficient
    cube (from _band_modes) on the n^3 grid of the unit box, [k][j][i]
    layout.  stagger[a] = 0 samples axis a at the cell centres (i+1/2)/n,
    1 at the lower faces i/n.  Evaluated as three separable sums, one axis
    at a time (the band has few modes), instead of an n^3 inverse FFT."""
    K = (coef.shape[0] - 1) // 2
    k1 = np.arange(-K, K + 1, dtype=np.float64)
    ex = [np.exp(2j * math.pi * np.outer((np.arange(n) + (0.0 if stagger[a] else 0.5)) / n, k1))
          for a in range(3)]                                  # [position, k] per axis
    t = np.einsum("zyk,xk->zyx", coef, ex[0])                 # (kz, ky, x)
    t = np.einsum("zkx,yk->zyx", t, ex[1])                    # (kz, y, x)
    out = np.empty((n, n, n))
    flat = t.reshape(t.shape[0], n * n)
    for z0 in range(0, n, 32):                                # slabs of z: bounded memory
        z1 = min(n, z0 + 32)
        out[z0:z1] = (ex[2][z0:z1] @ flat).real.reshape(z1 - z0, n, n)
    return out


def turbulence(n, ng=3, seed_b=1001, seed_v=1002, b0=0.2, db_rms=0.05, mach=0.5,
               rho0=1.0, p0=0.6, kmin=2.0, kmax=8.0):
    """3D decaying MHD turbulence on the periodic unit cube.

    rho=1, p=0.6 (c_s=1 at gamma=5/3), uniform Bx=b0.  The velocity is a
    solenoidal Gaussian field with a k^-2 shell spectrum on kmin<=|k|<=kmax
    (|v_k| ~ k^-2, coefficients projected perpendicular to k) and rms |v| =
    mach, sampled at the cell centres.  The field perturbation is the
    discrete curl of an edge-centred Gaussian vector potential with |A_k| ~
    k^-3 (k^-2 shell spectrum of B) scaled to rms |dB| = db_rms; it is
    solenoidal to round-off on the staggered grid.  The random coefficients
    are drawn per mode of the band from PCG64(seed), so the realisation is
    the same function of position at every n."""
    s = GridShape(n, n, n, ng)
    prim, bfld = s.empty()
    d = 1.0 / n
    rng = np.random.Generator(np.random.PCG64(seed_v))
    vk, (KX, KY, KZ) = _band_modes(rng, -2.0, kmin, kmax, 3)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[k2 == 0] = 1.0
    kdotv = KX * vk[0] + KY * vk[1] + KZ * vk[2]
    vk[0] -= KX * kdotv / k2
    vk[1] -= KY * kdotv / k2
    vk[2] -= KZ * kdotv / k2
    v = [_synth(vk[c], n, (0, 0, 0)) for c in range(3)]
    vscale = mach / math.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    # vector potential on cell edges: A_x at (i, j-1/2, k-1/2),
    # A_y at (i-1/2, j, k-1/2), A_z at (i-1/2, j-1/2, k)
    rng = np.random.Generator(np.random.PCG64(seed_b))
    ak, _ = _band_modes(rng, -3.0, kmin, kmax, 3)
    ax = _synth(ak[0], n, (0, 1, 1))
    ay = _synth(ak[1], n, (1, 0, 1))
    az = _synth(ak[2], n, (1, 1, 0))

    def curl(ax, ay, az):
        # arrays are [k][j][i]; roll(-1) along an axis reads index+1
        bx = (np.roll(az, -1, axis=1) - az) / d - (np.roll(ay, -1, axis=0) - ay) / d
        by = (np.roll(ax, -1, axis=0) - ax) / d - (np.roll(az, -1, axis=2) - az) / d
        bz = (np.roll(ay, -1, axis=2) - ay) / d - (np.roll(ax, -1, axis=1) - ax) / d
        return bx, by, bz

    bx, by, bz = curl(ax, ay, az)
    # face-to-centre average for the rms of the perturbation
    msq = np.mean((0.5 * (bx + np.roll(bx, -1, axis=2))) ** 2)
    msq += np.mean((0.5 * (by + np.roll(by, -1, axis=1))) ** 2)
    msq += np.mean((0.5 * (bz + np.roll(bz, -1, axis=0))) ** 2)
    scale = db_rms / math.sqrt(msq)
    del bx, by, bz
    bx, by, bz = curl(ax * scale, ay * scale, az * scale)
    del ax, ay, az
    _put(prim[RHO], s, rho0)
    _put(prim[PR], s, p0)
    _put(prim[VX], s, v[0] * vscale)
    _put(prim[VY], s, v[1] * vscale)
    _put(prim[VZ], s, v[2] * vscale)
    del v
    _put(bfld[0], s, bx + b0)
    _put(bfld[1], s, by)
    _put(bfld[2], s, bz)
    return s, prim, bfld


def brio_wu(n, axis=0, ng=3, transverse=4):
    """Brio-Wu shock tube (src/problem/st1d.pyx:66-83) laid along ``axis``
    of a grid with ``transverse`` cells in the other active axes (1D if
    transverse is 0).  Left: rho=1, p=1, B=(0.75, 1, 0); right: rho=0.125,
    p=0.1, B=(0.75, -1, 0), rotated so that the normal component lies along
    ``axis``; the interface is at the box centre."""
    sizes = [transverse] * 3 if transverse else [1, 1, 1]
    sizes[axis] = n
    if not transverse:
        if axis != 0:
            raise ValueError("1D tubes lie along x")
    s = GridShape(sizes[0], sizes[1], sizes[2], ng)
    prim, bfld = s.empty()
    xc = s.centres(axis)
    left = xc <= 0.5
    shape = [1, 1, 1]
    shape[2 - axis] = n
    L = left.reshape(shape)
    sk, sj, si = s.active()
    act = (slice(None), sk, sj, si)
    full = np.ones((s.nz if s.dims >= 3 else 1, s.ny if s.dims >= 2 else 1, s.nx))
    prim[RHO][act[1:]] = np.where(L, 1.0, 0.125) * full
    prim[PR][act[1:]] = np.where(L, 1.0, 0.1) * full
    bn, bt = 0.75, np.where(L, 1.0, -1.0) * full
    # components (normal, first transverse, second transverse) in cyclic order
    tr1 = (axis + 1) % 3
    bfld[axis][act[1:]] = bn * full
    bfld[tr1][act[1:]] = bt
    return s, prim, bfld


def sod(n, axis=0, ng=3, transverse=0, vel=0.0):
    """Sod's shock tube for the hydro build (no field): left rho=1, p=1,
    right rho=0.125, p=0.1, both moving with ``vel`` along the tube, laid
    along ``axis`` with ``transverse`` cells in the other active axes (1D
    if 0); interface at the box centre."""
    sizes = [transverse] * 3 if transverse else [1, 1, 1]
    sizes[axis] = n
    if not transverse and axis != 0:
        raise ValueError("1D tubes lie along x")
    s = GridShape(sizes[0], sizes[1], sizes[2], ng)
    prim, bfld = s.empty()
    shape = [1, 1, 1]
    shape[2 - axis] = n
    L = (s.centres(axis) <= 0.5).reshape(shape)
    full = np.ones((s.nz if s.dims >= 3 else 1, s.ny if s.dims >= 2 else 1, s.nx))
    act = s.active()
    prim[RHO][act] = np.where(L, 1.0, 0.125) * full
    prim[PR][act] = np.where(L, 1.0, 0.1) * full
    prim[VX + axis][act] = vel * full
    return s, prim, bfld


def hydro_blast(n, ng=3, dims=2, p_in=10.0, p_out=0.1, radius=0.1, seed=7):
    """Blast without field for the hydro build: rho=1 (seeded +-5 %
    perturbation), p=p_out, p=p_in inside r<radius of the box centre, on a
    periodic square (dims=2) or cube (dims=3)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = GridShape(n, n, n if dims == 3 else 1, ng)
    prim, bfld = s.empty()
    X = np.meshgrid(*[s.centres(a) for a in range(dims)][::-1], indexing="ij")
    r2 = sum((x - 0.5) ** 2 for x in X)
    shp = prim[RHO][s.active()].shape
    _put(prim[RHO], s, (1.0 + 0.05 * rng.uniform(-1, 1, shp)))
    _put(prim[PR], s, np.where(r2.reshape(shp) < radius ** 2, p_in, p_out))
    _put(prim[PSC], s, np.where(r2.reshape(shp) < radius ** 2, 1.0, 0.0))
    return s, prim, bfld


def linear_wave(n, wave="fast", amp=1e-6, ng=3):
    """1D linear MHD eigenmode on [0,1] (state of src/problem/linwaves1d.pyx:
    48-104): rho=1, p=0.6, B=(1, sqrt 2, 1/2), gamma=5/3 (c_s=1, c_A,x=1,
    c_fast=2, c_slow=1/2), perturbation amp*cos(2 pi x) times the eigenvector
    of the fast, slow or Alfven wave.  The pressure perturbation is the
    adiabatic one, dp = c_s^2 drho; the reference's commented formulas put the
    conserved energy perturbation of the eigenvector into p (linwaves1d.pyx:
    91,103), which is not an eigenmode."""
    s = GridShape(n, 1, 1, ng)
    prim, bfld = s.empty()
    x = s.centres(0)
    a = amp * np.cos(2 * math.pi * x)
    r5 = math.sqrt(5.0)
    if wave == "fast":
        ev = dict(rho=1 / r5, vx=-2 / r5, vy=2 * math.sqrt(2) / 3 / r5, vz=1.0 / 3 / r5,
                  by=4 * math.sqrt(2) / 3 / r5, bz=2.0 / 3 / r5, p=1.0 / r5)
    elif wave == "slow":
        ev = dict(rho=2 / r5, vx=-1 / r5, vy=-4 * math.sqrt(2) / 3 / r5, vz=-2.0 / 3 / r5,
                  by=-2 * math.sqrt(2) / 3 / r5, bz=-1.0 / 3 / r5, p=2.0 / r5)
    elif wave == "alfven":
        ev = dict(rho=0.0, vx=0.0, vy=-1.0 / 3, vz=2 * math.sqrt(2) / 3, by=-1.0 / 3,
                  bz=2 * math.sqrt(2) / 3, p=0.0)
    else:
        raise ValueError(wave)
    _put(prim[RHO], s, 1.0 + ev["rho"] * a)
    _put(prim[VX], s, ev["vx"] * a)
    _put(prim[VY], s, ev["vy"] * a)
    _put(prim[VZ], s, ev["vz"] * a)
    _put(prim[PR], s, 0.6 + ev["p"] * a)
    _put(bfld[0], s, 1.0)
    _put(bfld[1], s, math.sqrt(2.0) + ev["by"] * a)
    _put(bfld[2], s, 0.5 + ev["bz"] * a)
    return s, prim, bfld


# Eigenvectors of the linear waves of the state rho=1, p=0.6, B=(1, sqrt 2,
# 1/2), gamma=5/3 in the frame (normal, first transverse, second transverse),
# per unit amplitude; every mode travels towards -normal at its speed.
_R5 = math.sqrt(5.0)
WAVES = {
    "fast": (2.0, dict(rho=1 / _R5, vn=-2 / _R5, v1=2 * math.sqrt(2) / 3 / _R5, v2=1.0 / 3 / _R5,
                       b1=4 * math.sqrt(2) / 3 / _R5, b2=2.0 / 3 / _R5, p=1.0 / _R5)),
    "alfven": (1.0, dict(rho=0.0, vn=0.0, v1=-1.0 / 3, v2=2 * math.sqrt(2) / 3, b1=-1.0 / 3,
                         b2=2 * math.sqrt(2) / 3, p=0.0)),
    "slow": (0.5, dict(rho=2 / _R5, vn=-1 / _R5, v1=-4 * math.sqrt(2) / 3 / _R5,
                       v2=-2.0 / 3 / _R5, b1=-2 * math.sqrt(2) / 3 / _R5, b2=-1.0 / 3 / _R5,
                       p=2.0 / _R5)),
}


def oblique_frame(dims):
    """Unit vectors (normal, e1, e2) of a right-handed frame whose normal is
    the diagonal of the unit box in ``dims`` dimensions: (1,0,0) in 1D,
    (1,1,0)/sqrt 2 in 2D, (1,1,1)/sqrt 3 in 3D, with e2 = normal x e1."""
    if dims == 1:
        kn = np.array([1.0, 0.0, 0.0])
        e1 = np.array([0.0, 1.0, 0.0])
    elif dims == 2:
        kn = np.array([1.0, 1.0, 0.0]) / math.sqrt(2.0)
        e1 = np.array([-1.0, 1.0, 0.0]) / math.sqrt(2.0)
    else:
        kn = np.array([1.0, 1.0, 1.0]) / math.sqrt(3.0)
        e1 = np.array([-1.0, 1.0, 0.0]) / math.sqrt(2.0)
    return kn, e1, np.cross(kn, e1)


def linear_wave_oblique(n, wave="fast", dims=2, amp=1e-6, t=0.0, ng=3):
    """Linear MHD eigenmode (the waves of linear_wave) propagating along the
    diagonal of the periodic unit box in ``dims`` dimensions (n cells per
    axis), with wave vector k = 2 pi (1, 1, 1)[:dims], so one wavelength is
    1/sqrt(dims) along the diagonal; the state at time ``t`` of the exact
    linear solution, amp * cos(k.x + omega t) times the eigenvector rotated
    onto oblique_frame(dims).  Cell variables are point values at the cell
    centres; the background field sits on the faces as it is, and the
    perturbation of the field is the discrete curl of an edge-centred
    vector potential, so the face field is solenoidal to round-off.
    Returns (shape, prim, bfld, period)."""
    speed, ev = WAVES[wave]
    kn, e1, e2 = oblique_frame(dims)
    kvec = 2 * math.pi * np.array([1.0 if a < dims else 0.0 for a in range(3)])
    kmag = float(np.linalg.norm(kvec))
    omega = speed * kmag
    s = GridShape(n, n if dims >= 2 else 1, n if dims >= 3 else 1, ng)
    prim, bfld = s.empty()
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h  # centres
    f = np.arange(n) * h          # lower faces
    one = np.zeros(1)

    def axes(xs, ys, zs):  # arrays [k][j][i]
        Z, Y, X = np.meshgrid(zs if dims >= 3 else one, ys if dims >= 2 else one, xs,
                              indexing="ij")
        return X, Y, Z

    def phase(X, Y, Z):
        return kvec[0] * X + kvec[1] * Y + kvec[2] * Z + omega * t

    X, Y, Z = axes(c, c, c)
    a = amp * np.cos(phase(X, Y, Z))
    b0 = 1.0 * kn + math.sqrt(2.0) * e1 + 0.5 * e2
    dv = ev["vn"] * kn + ev["v1"] * e1 + ev["v2"] * e2
    db = ev["b1"] * e1 + ev["b2"] * e2
    _put(prim[RHO], s, 1.0 + ev["rho"] * a)
    _put(prim[PR], s, 0.6 + ev["p"] * a)
    for q, var in enumerate((VX, VY, VZ)):
        _put(prim[var], s, dv[q] * a)
    # perturbation of B: curl of A = Ahat sin(k.x + omega t) is
    # cos(.) k x Ahat, which equals db cos(.) for Ahat = -(k x db) / |k|^2
    ahat = -amp * np.cross(kvec, db) / kmag ** 2
    # edges: A_x at (centre x, face y, face z), A_y at (face x, centre y, face z),
    # A_z at (face x, face y, centre z); absent axes sit at 0
    ax = ahat[0] * np.sin(phase(*axes(c, f, f)))
    ay = ahat[1] * np.sin(phase(*axes(f, c, f)))
    az = ahat[2] * np.sin(phase(*axes(f, f, c)))

    def up(arr, axis):  # value at index + 1 (periodic) along grid axis 0=x
        if axis >= dims:
            return arr
        return np.roll(arr, -1, axis=2 - axis)

    bx = (up(az, 1) - az) / h - (up(ay, 2) - ay) / h
    by = (up(ax, 2) - ax) / h - (up(az, 0) - az) / h
    bz = (up(ay, 0) - ay) / h - (up(ax, 1) - ax) / h
    # components along absent axes: the face value is the point value there
    if dims < 2:
        by = db[1] * a
    if dims < 3:
        bz = db[2] * a
    _put(bfld[0], s, b0[0] + bx)
    _put(bfld[1], s, b0[1] + by)
    _put(bfld[2], s, b0[2] + bz)
    return s, prim, bfld, 1.0 / math.sqrt(dims) / speed


def blast2d(n, ng=3, seed=5005, p_in=10.0, p_out=0.1, radius=0.1, theta=None, bmag=None):
    """2D MHD blast on the periodic unit square: rho=1, p=p_out, uniform B at
    a random angle with |B| in [0.5, 2], p=p_in inside r<radius of the centre
    (cf. src/problem/blast_MHD.pyx:24-59).  theta and bmag, when given,
    replace the seeded draws."""
    rng = np.random.Generator(np.random.PCG64(seed))
    th = rng.uniform(0.0, 2 * math.pi)
    bm = rng.uniform(0.5, 2.0)
    theta = th if theta is None else theta
    bmag = bm if bmag is None else bmag
    s = GridShape(n, n, 1, ng)
    prim, bfld = s.empty()
    x = s.centres(0)
    y = s.centres(1)
    X, Y = np.meshgrid(x, y, indexing="xy")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    _put(prim[RHO], s, 1.0)
    _put(prim[PR], s, np.where(r < radius, p_in, p_out))
    _put(bfld[0], s, bmag * math.cos(theta))
    _put(bfld[1], s, bmag * math.sin(theta))
    return s, prim, bfld


def _rot12(a):
    """Tube frame -> grid frame of the oblique tube: the tube's x axis along
    (1, 2)/sqrt 5 (st2d.pyx:23-30 is the same rotation)."""
    r5 = math.sqrt(5.0)
    return np.array([(a[0] - 2.0 * a[1]) / r5, (2.0 * a[0] + a[1]) / r5, a[2]])


def oblique_tube(ny, ng=3, aspect=4):
    """The reference's 2D oblique shock tube (src/problem/st2d.pyx): the
    Ryu-Jones 2a Riemann problem (left rho=1.08, v=(1.2, 0.01, 0.5), p=0.95,
    B=(2, 3.6, 2)/sqrt(4 pi); right rho=1, v=0, p=1, B=(2, 4, 2)/sqrt(4 pi))
    with its normal along (1, 2)/sqrt 5 on an (aspect ny) x ny grid of
    [0, aspect] x [0, 1] (square cells), the interface through the box
    centre.  With aspect 4 the interface stays ny cells away from both x
    sides, whose fixed states are the far-field ones.  Every cell, ghosts
    included, is set from the integer test i' + 2 j' <= nx/2 + ny - 2
    (i', j' counted from the first active cell), so the state is exactly
    invariant under (i, j) -> (i - 2, j + 1).  Face fields are differences of
    the corner potential A_z = J' dx (Bx + 2 By) + g(I' + 2 J'), with g
    continuous and linear on each side (g' = -By dx per unit), which curls to
    the two states (Bx + 2 By = sqrt 5 B . n is the same on both); each face
    value is computed from I' + 2 J' alone, so the face fields are exactly
    invariant too.

    Returns (shape, prim, bfld, user): user maps (axis, side) to the user
    boundaries of st2d.pyx:198-357: fixed left and right states on the x
    sides over the active rows, and on the y sides the opposite active rows
    shifted by -+2 ny cells along x.  st2d clamps the shifted column to the
    grid (:318-320, 344-346), which reads the x sides' ghost cells before
    the x sides fill them (the y sides come first), that is, whatever those
    cells held between fills; here the clamp is to the first and last active
    columns, which hold the same far-field states."""
    nx = aspect * ny
    if nx % 2:
        raise ValueError("oblique_tube: aspect * ny must be even")
    s = GridShape(nx, ny, 1, ng, (0.0, 0.0, 0.0), (float(aspect), 1.0, 1.0))
    prim, bfld = s.empty()
    nt = s.ntot
    c4 = 1.0 / math.sqrt(4.0 * math.pi)
    sides = []
    for rho, v, p, b in ((1.08, (1.2, 0.01, 0.5), 0.95, (2.0, 3.6, 2.0)),
                         (1.0, (0.0, 0.0, 0.0), 1.0, (2.0, 4.0, 2.0))):
        w = np.zeros(9)
        w[RHO], w[PR] = rho, p
        w[VX:VZ + 1] = _rot12(v)
        w[BXC:BZC + 1] = _rot12(np.array(b) * c4)
        sides.append(w)
    ip = np.arange(nt[0]) - ng
    jp = np.arange(nt[1]) - ng
    left = (ip[None, :] + 2 * jp[:, None]) <= nx // 2 + ny - 2
    for n in (RHO, VX, VY, VZ, PR):
        prim[n, 0] = np.where(left, sides[0][n], sides[1][n])
    dx = 1.0 / ny
    u0 = nx // 2 + ny  # corners with I' + 2 J' < u0 lie on the left
    kb = sides[0][BXC] + 2.0 * sides[0][BYC]
    byl, byr = sides[0][BYC], sides[1][BYC]

    def g(u):
        return np.where(u < u0, -byl * u * dx, -byr * u * dx + (byr - byl) * u0 * dx)

    # face (I, j) of bx spans corners u and u + 2, face (i, J) of by u and u + 1
    u = ip[None, :] + 2 * jp[:, None]
    bfld[0, 0] = kb + (g(u + 2) - g(u)) / dx
    bfld[1, 0] = -(g(u + 1) - g(u)) / dx
    bfld[2, 0] = sides[0][BZC]
    from .user_bc import Fixed, Mapped
    i1, i2 = ng, ng + nx - 1
    user = {
        (0, 0): Fixed(sides[0], rows="active"),
        (0, 1): Fixed(sides[1], rows="active"),
        (1, 0): Mapped(lambda i, j, k: (np.clip(i - 2 * ny, i1, i2), j + ny, k)),
        (1, 1): Mapped(lambda i, j, k: (np.clip(i + 2 * ny, i1, i2), j - ny, k)),
    }
    return s, prim, bfld, user


def riemann_tube(n, rng, ng=3):
    """One seeded 1D MHD Riemann problem on [0,1]: rho in [0.1,1], p in
    [0.1,1], |v| <= 1 along a random direction, Bx in [0,1] shared by both
    sides, By, Bz in [-1,1]; interface at x=0.5."""
    s = GridShape(n, 1, 1, ng)
    prim, bfld = s.empty()
    x = s.centres(0)
    left = x <= 0.5
    bx = rng.uniform(0.0, 1.0)
    states = []
    for _ in range(2):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        v = d * rng.uniform(0.0, 1.0)
        states.append(dict(rho=rng.uniform(0.1, 1.0), p=rng.uniform(0.1, 1.0), v=v,
                           by=rng.uniform(-1.0, 1.0), bz=rng.uniform(-1.0, 1.0)))
    L, R = states
    _put(prim[RHO], s, np.where(left, L["rho"], R["rho"]))
    _put(prim[PR], s, np.where(left, L["p"], R["p"]))
    for c, var in enumerate((VX, VY, VZ)):
        _put(prim[var], s, np.where(left, L["v"][c], R["v"][c]))
    _put(bfld[0], s, bx)
    _put(bfld[1], s, np.where(left, L["by"], R["by"]))
    _put(bfld[2], s, np.where(left, L["bz"], R["bz"]))
    return s, prim, bfld


def tube_sweep(nbox=512, seed=4004, nmin=64, nmax=4096, ng=3):
    """Batch of seeded 1D Riemann problems of uneven length: lengths drawn
    uniformly from [nmin, nmax], then one riemann_tube per box from the same
    generator.  Returns a list of (shape, prim, bfld)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = rng.integers(nmin, nmax + 1, size=nbox)
    return [riemann_tube(int(n), rng, ng) for n in lengths]


def blast_sweep(nbox=64, seed=5005, nmin=64, nmax=512, ng=3):
    """Batch of 2D periodic blast boxes of uneven side drawn uniformly from
    [nmin, nmax], each with its own field angle in [0, 2pi) and |B| in
    [0.5, 2].  Returns a list of (shape, prim, bfld)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = rng.integers(nmin, nmax + 1, size=nbox)
    out = []
    for n in sides:
        theta = rng.uniform(0.0, 2 * math.pi)
        bmag = rng.uniform(0.5, 2.0)
        out.append(blast2d(int(n), ng, theta=theta, bmag=bmag))
    return out
