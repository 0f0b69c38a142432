"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``).  Only the tests,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this
module; the product path in ``bonzo_amd`` never does.

Layout of every array is the reference's: C-contiguous ``[var][k][j][i]``
(src/coordinates/grid.pyx:28-52), ``prim``/``cons`` with 9 rows and ``bfld``
with 3 rows indexed by the lower face (src/integration/ct.pyx:110-111).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_libs = {}

RS_IDS = {"hll": 0, "hllc": 1, "hlld": 2, "hlla": 3, "hllt": 4}

_dp = ctypes.POINTER(ctypes.c_double)
_dpp = ctypes.POINTER(_dp)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def mutant_path(name):
    """Path of a mutant oracle (``make mutants``: one planted mistake, for the
    oracle's own sensitivity tests), built on first use."""
    path = os.path.join(_HERE, "_mut", f"liboracle_{name}.so")
    subprocess.run(["make", "-s", "-C", _HERE, os.path.join("_mut", f"liboracle_{name}.so")],
                   check=True)
    return path


def lib(path=None):
    path = path or _LIB_PATH
    if path not in _libs:
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        L.or_create_ex.restype = ctypes.c_void_p
        L.or_create_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
        L.or_rk3_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        for name in ("or_prim", "or_cons", "or_bfld"):
            getattr(L, name).restype = _dp
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.or_coord.restype = _dp
        L.or_coord.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_ntot.restype = ctypes.c_int
        L.or_ntot.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_ng.restype = ctypes.c_int
        L.or_ng.argtypes = [ctypes.c_void_p]
        L.or_set_dt_fms_ref.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_char_proj.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ("or_prim2char", "or_char2prim"):
            getattr(L, name).argtypes = [_dp, _dp, ctypes.c_double]
        L.or_apply_grid_bc.argtypes = [ctypes.c_void_p]
        L.or_init.argtypes = [ctypes.c_void_p]
        L.or_new_dt.restype = ctypes.c_double
        L.or_new_dt.argtypes = [ctypes.c_void_p]
        L.or_vl_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_advance.restype = ctypes.c_int
        L.or_advance.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.or_advance_steps.restype = ctypes.c_int
        L.or_advance_steps.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_time.restype = ctypes.c_double
        L.or_time.argtypes = [ctypes.c_void_p]
        L.or_step.restype = ctypes.c_int
        L.or_step.argtypes = [ctypes.c_void_p]
        L.or_last_dt.restype = ctypes.c_double
        L.or_last_dt.argtypes = [ctypes.c_void_p]
        L.or_history.argtypes = [ctypes.c_void_p, _dp]
        L.or_history.restype = None
        for name in ("or_cons2prim", "or_prim2cons"):
            getattr(L, name).argtypes = [_dp, _dp, ctypes.c_double]
        L.or_primcons2flux.argtypes = [_dp, _dp, _dp, ctypes.c_double]
        L.or_fms.restype = ctypes.c_double
        L.or_fms.argtypes = [_dp, ctypes.c_double, ctypes.c_double]
        L.or_vl_lim.restype = ctypes.c_double
        L.or_vl_lim.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_fms_hydro.restype = ctypes.c_double
        L.or_fms_hydro.argtypes = [_dp, ctypes.c_double]
        L.or_set_hydro.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_ct_upwind.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_user_bc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), _dp]
        for name in ("or_flux",):
            getattr(L, name).restype = _dp
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_edge_emf.restype = _dp
        L.or_edge_emf.argtypes = [ctypes.c_void_p]
        L.or_corner_emf.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        for name in ("or_hll_flux", "or_hllt_flux", "or_hlld_flux", "or_hll_flux_hydro",
                     "or_hllc_flux"):
            getattr(L, name).argtypes = [_dpp, _dpp, _dpp, _dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double]
        for name in ("or_reconstr_const", "or_reconstr_linear"):
            getattr(L, name).argtypes = [_dpp, _dpp, ctypes.POINTER(_dpp),
                                         ctypes.POINTER(_dpp), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]
        _libs[path] = L
    return _libs[path]


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _rows(a):
    """double** over the rows of a C-contiguous 2D float64 array."""
    arr = (_dp * a.shape[0])(*[ctypes.cast(a[n].ctypes.data, _dp) for n in range(a.shape[0])])
    return arr


# ---------------------------------------------------------------- point API
def fms(w, gam, bx=None):
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().or_fms(_ptr(w), float(w[6] if bx is None else bx), gam)


def prim2cons(w, gam):
    w = np.ascontiguousarray(w, dtype=np.float64)
    u = np.zeros(9)
    lib().or_prim2cons(_ptr(u), _ptr(w), gam)
    return u


def cons2prim(u, gam):
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.zeros(9)
    lib().or_cons2prim(_ptr(w), _ptr(u), gam)
    return w


def primcons2flux(w, gam):
    w = np.ascontiguousarray(w, dtype=np.float64)
    u = prim2cons(w, gam)
    f = np.zeros(9)
    lib().or_primcons2flux(_ptr(f), _ptr(w), _ptr(u), gam)
    return f


def vl_lim(a, b):
    return lib().or_vl_lim(float(a), float(b))


def prim2char(vc, w, gam):
    """Characteristic amplitudes of the primitive vector vc in the eigensystem
    of state w (src/mhd/eos_c.h:100-186)."""
    vc = np.array(vc, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().or_prim2char(_ptr(vc), _ptr(w), gam)
    return vc


def char2prim(vc, w, gam):
    """Inverse of prim2char (src/mhd/eos_c.h:191-280)."""
    vc = np.array(vc, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().or_char2prim(_ptr(vc), _ptr(w), gam)
    return vc


def riemann(rsolver, wl, wr, bx, gam):
    """Pencil call of a solver: wl, wr are [9, n] primitive states, bx [n].

    Returns the [9, n] flux.  Same seam as src/mhd/fluxes_c.h:3-21."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    bx = np.ascontiguousarray(bx, dtype=np.float64)
    n = wl.shape[1]
    flux = np.full((9, n), np.nan)
    fn = {"hll": lib().or_hll_flux, "hlld": lib().or_hlld_flux, "hllt": lib().or_hllt_flux,
          "hll_hydro": lib().or_hll_flux_hydro, "hllc": lib().or_hllc_flux}[rsolver]
    fn(_rows(flux), _rows(wl), _rows(wr), _ptr(bx), 0, n - 1, gam)
    return flux


def reconstr_x(w, order):
    """x-pencil reconstruction of a [9, n] row: returns (wl, wr) as the
    reference's scratch rows (wl[i+1], wr[i] from cell i), src/coordinates/
    reconstr_c.c:8-91, over cells 1..n-2."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = w.shape[1]
    pad = 4
    wl = np.zeros((9, n + pad))
    wr = np.zeros((9, n + pad))
    scr = [np.zeros((9, n + pad)) for _ in range(3)]
    keep = []
    stencil = []
    for d in range(2 * order + 1):
        rows = (_dp * 9)()
        for v in range(9):
            rows[v] = ctypes.cast(w[v].ctypes.data + 8 * (d - order), _dp)
        keep.append(rows)
        stencil.append(ctypes.cast(rows, _dpp))
    st = (_dpp * 3)(*stencil + [stencil[0]] * (3 - len(stencil)))
    scr_rows = [_rows(a) for a in scr]
    sc = (_dpp * 3)(*[ctypes.cast(r, _dpp) for r in scr_rows])
    fn = lib().or_reconstr_linear if order == 1 else lib().or_reconstr_const
    fn(_rows(wl), _rows(wr), st, sc, 0, order, n - 1 - order)
    return wl[:, :n], wr[:, :n]


# ---------------------------------------------------------------- grid API
class OracleSim:
    """Whole-grid oracle run with the reference's VL sequencing."""

    def __init__(self, nx, ny=1, nz=1, reconstr_order=1, lmin=(0., 0., 0.),
                 lmax=(1., 1., 1.), gam=5. / 3., cour=0.4, rsolver="hlld",
                 bc=(0, 0, 0, 0, 0, 0), nthreads=None, dt_fms_ref=True, lib_path=None,
                 char_proj=False, hydro=False, ct_upwind=False, tintegr="vl"):
        L = self._L = lib(lib_path)
        lmin = np.asarray(lmin, dtype=np.float64)
        lmax = np.asarray(lmax, dtype=np.float64)
        bcs = (ctypes.c_int * 6)(*bc)
        nthreads = nthreads or min(8, os.cpu_count() or 1)
        self._h = L.or_create_ex(nx, ny, nz, reconstr_order, _ptr(lmin), _ptr(lmax), gam, cour,
                                 RS_IDS[rsolver], bcs, nthreads, 1 if tintegr == "rk3" else 0)
        L.or_set_dt_fms_ref(self._h, int(bool(dt_fms_ref)))
        L.or_set_char_proj(self._h, int(bool(char_proj)))
        L.or_set_hydro(self._h, int(bool(hydro)))
        L.or_set_ct_upwind(self._h, int(bool(ct_upwind)))
        self.ntot = tuple(L.or_ntot(self._h, a) for a in range(3))
        self.ng = L.or_ng(self._h)
        nc = self.ntot[0] * self.ntot[1] * self.ntot[2]
        shp = (self.ntot[2], self.ntot[1], self.ntot[0])
        self.prim = np.ctypeslib.as_array(L.or_prim(self._h), shape=(9 * nc,)).reshape((9,) + shp)
        self.cons = np.ctypeslib.as_array(L.or_cons(self._h), shape=(9 * nc,)).reshape((9,) + shp)
        self.bfld = np.ctypeslib.as_array(L.or_bfld(self._h), shape=(3 * nc,)).reshape((3,) + shp)

    def set_user_bc(self, axis, side, src, fixed=None):
        """User boundary (flag 3) of one side, raw arrays in the layout of
        bnz_set_user_bc (src per slab cell, fixed [n][12])."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        fp = None
        if fixed is not None:
            fixed = np.ascontiguousarray(fixed, dtype=np.float64)
            fp = _ptr(fixed)
        rc = self._L.or_set_user_bc(self._h, axis, side,
                                    src.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), fp)
        if rc != 0:
            raise ValueError("or_set_user_bc: bad side or source index")

    def coord(self, which, axis):
        n = self.ntot[axis] + (1 if which == 0 else 0)
        return np.ctypeslib.as_array(self._L.or_coord(self._h, which, axis), shape=(n,)).copy()

    def flux(self, axis):
        """The face-flux array of one axis, [9][k][j][i] (a view)."""
        nc = self.ntot[0] * self.ntot[1] * self.ntot[2]
        shp = (9, self.ntot[2], self.ntot[1], self.ntot[0])
        return np.ctypeslib.as_array(self._L.or_flux(self._h, axis), shape=(9 * nc,)).reshape(shp)

    def corner_emf(self, upwind, dt):
        """One corner-EMF evaluation from the current prim and flux arrays;
        returns a copy of the edge EMFs [3][k][j][i]."""
        self._L.or_corner_emf(self._h, int(bool(upwind)), dt)
        nc = self.ntot[0] * self.ntot[1] * self.ntot[2]
        shp = (3, self.ntot[2], self.ntot[1], self.ntot[0])
        return np.ctypeslib.as_array(self._L.or_edge_emf(self._h), shape=(3 * nc,)).reshape(
            shp).copy()

    def init(self):
        self._L.or_init(self._h)

    def apply_bc(self):
        self._L.or_apply_grid_bc(self._h)

    def new_dt(self):
        return self._L.or_new_dt(self._h)

    def vl_step(self, dt):
        self._L.or_vl_step(self._h, dt)

    def rk3_step(self, dt):
        self._L.or_rk3_step(self._h, dt)

    def advance(self, tmax, max_steps=0):
        return self._L.or_advance(self._h, tmax, max_steps)

    def advance_steps(self, n):
        return self._L.or_advance_steps(self._h, n)

    @property
    def time(self):
        return self._L.or_time(self._h)

    @property
    def step(self):
        return self._L.or_step(self._h)

    @property
    def last_dt(self):
        return self._L.or_last_dt(self._h)

    def history(self):
        out = np.zeros(16)
        self._L.or_history(self._h, _ptr(out))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._L.or_destroy(self._h)
            self._h = None
            self.prim = self.cons = self.bfld = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
