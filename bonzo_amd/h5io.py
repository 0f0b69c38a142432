"""HDF5 grid and restart files in the reference's format, written and read
through the HDF5 C library (libhdf5, loaded with ctypes; h5py, which the
reference uses, is not required).

Layout (src/io/output.pyx:248-385 write_grid, src/io/restart.pyx:23-86):
  datasets  rho, vx, vy, vz, p, psc, bxc, byc, bzc, bxf, byf, bzf (primitive
            variables; the conserved set writes rho, mx, my, mz, en, psc in
            the first six) as float64 [Nz][Ny][Nx] over the active cells (or
            [Ntot_z][Ntot_y][Ntot_x] with ghosts), chunked as one block per
            dataset when ghosts are left out, as the reference's single-block
            write is;
  attributes time (float64), step (int64), dt (float64), xf, yf, zf (the
            face coordinates, float64 arrays of Ntot + 1).
File names: out/grid_NNN.h5 with NNN = int(time / grid_dt), rst/rst_grid.h5.
"""
import ctypes
import glob
import os
import sys

import numpy as np

PRIM_NAMES = ("rho", "vx", "vy", "vz", "p", "psc", "bxc", "byc", "bzc")
CONS_NAMES = ("rho", "mx", "my", "mz", "en", "psc", "bxc", "byc", "bzc")
FACE_NAMES = ("bxf", "byf", "bzf")

_H5F_ACC_RDONLY, _H5F_ACC_TRUNC = 0, 2
_H5P_DEFAULT, _H5S_ALL, _H5S_SCALAR = 0, 0, 0
_hid = ctypes.c_int64
_hsize = ctypes.c_ulonglong
_L = None


def _find_library():
    cands = []
    env = os.environ.get("BNZ_HDF5_LIB")
    if env:
        cands.append(env)
    for d in (os.path.join(sys.prefix, "lib"), "/opt/conda/lib", "/usr/lib/x86_64-linux-gnu",
              "/usr/local/lib"):
        cands += sorted(glob.glob(os.path.join(d, "libhdf5.so*")))
    for c in cands:
        base = os.path.basename(c)
        if base.startswith("libhdf5.so") and os.path.exists(c):
            return c
    raise RuntimeError("the HDF5 C library (libhdf5.so) was not found; set BNZ_HDF5_LIB")


def lib():
    """The HDF5 C library, with the signatures used here."""
    global _L
    if _L is not None:
        return _L
    L = ctypes.CDLL(_find_library())
    vp, cp, hp = ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(_hsize)
    sigs = {
        "H5open": ([], ctypes.c_int),
        "H5Fcreate": ([cp, ctypes.c_uint, _hid, _hid], _hid),
        "H5Fopen": ([cp, ctypes.c_uint, _hid], _hid),
        "H5Fclose": ([_hid], ctypes.c_int),
        "H5Screate_simple": ([ctypes.c_int, hp, hp], _hid),
        "H5Screate": ([ctypes.c_int], _hid),
        "H5Sclose": ([_hid], ctypes.c_int),
        "H5Sget_simple_extent_ndims": ([_hid], ctypes.c_int),
        "H5Sget_simple_extent_dims": ([_hid, hp, hp], ctypes.c_int),
        "H5Pcreate": ([_hid], _hid),
        "H5Pclose": ([_hid], ctypes.c_int),
        "H5Pset_chunk": ([_hid, ctypes.c_int, hp], ctypes.c_int),
        "H5Pget_chunk": ([_hid, ctypes.c_int, hp], ctypes.c_int),
        "H5Pget_layout": ([_hid], ctypes.c_int),
        "H5Dcreate2": ([_hid, cp, _hid, _hid, _hid, _hid, _hid], _hid),
        "H5Dopen2": ([_hid, cp, _hid], _hid),
        "H5Dwrite": ([_hid, _hid, _hid, _hid, _hid, vp], ctypes.c_int),
        "H5Dread": ([_hid, _hid, _hid, _hid, _hid, vp], ctypes.c_int),
        "H5Dget_space": ([_hid], _hid),
        "H5Dget_type": ([_hid], _hid),
        "H5Dget_create_plist": ([_hid], _hid),
        "H5Dclose": ([_hid], ctypes.c_int),
        "H5Acreate2": ([_hid, cp, _hid, _hid, _hid, _hid], _hid),
        "H5Aopen": ([_hid, cp, _hid], _hid),
        "H5Awrite": ([_hid, _hid, vp], ctypes.c_int),
        "H5Aread": ([_hid, _hid, vp], ctypes.c_int),
        "H5Aget_space": ([_hid], _hid),
        "H5Aget_type": ([_hid], _hid),
        "H5Aclose": ([_hid], ctypes.c_int),
        "H5Aexists": ([_hid, cp], ctypes.c_int),
        "H5Lexists": ([_hid, cp, _hid], ctypes.c_int),
        "H5Tequal": ([_hid, _hid], ctypes.c_int),
        "H5Tclose": ([_hid], ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    _L = L
    return L


def _g(name):
    """Value of one of the library's global type or class identifiers."""
    return _hid.in_dll(lib(), name).value


def _ok(rc, what):
    if rc < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return rc


def _dims(shape):
    return (_hsize * len(shape))(*shape)


def write_h5(path, datasets, attrs, chunked=True):
    """Write float64 datasets (name -> C-contiguous 3D array) and attributes
    (name -> float, int or 1D float array) to a new HDF5 file at ``path``."""
    L = lib()
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    f = _ok(L.H5Fcreate(path.encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), "H5Fcreate")
    try:
        for name, arr in datasets.items():
            a = np.ascontiguousarray(arr, dtype=np.float64)
            sp = _ok(L.H5Screate_simple(a.ndim, _dims(a.shape), None), "H5Screate_simple")
            dcpl = _H5P_DEFAULT
            if chunked:
                dcpl = _ok(L.H5Pcreate(_g("H5P_CLS_DATASET_CREATE_ID_g")), "H5Pcreate")
                _ok(L.H5Pset_chunk(dcpl, a.ndim, _dims(a.shape)), "H5Pset_chunk")
            ds = _ok(L.H5Dcreate2(f, name.encode(), _g("H5T_IEEE_F64LE_g"), sp, _H5P_DEFAULT,
                                  dcpl, _H5P_DEFAULT), f"H5Dcreate2 {name}")
            _ok(L.H5Dwrite(ds, _g("H5T_NATIVE_DOUBLE_g"), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                           a.ctypes.data), f"H5Dwrite {name}")
            L.H5Dclose(ds)
            if chunked:
                L.H5Pclose(dcpl)
            L.H5Sclose(sp)
        for name, val in attrs.items():
            if isinstance(val, (int, np.integer)):
                buf = np.array(int(val), dtype=np.int64)
                ftype, mtype = _g("H5T_STD_I64LE_g"), _g("H5T_NATIVE_LLONG_g")
            else:
                buf = np.ascontiguousarray(val, dtype=np.float64)
                ftype, mtype = _g("H5T_IEEE_F64LE_g"), _g("H5T_NATIVE_DOUBLE_g")
            sp = (_ok(L.H5Screate(_H5S_SCALAR), "H5Screate") if buf.ndim == 0 else
                  _ok(L.H5Screate_simple(1, _dims(buf.shape), None), "H5Screate_simple"))
            at = _ok(L.H5Acreate2(f, name.encode(), ftype, sp, _H5P_DEFAULT, _H5P_DEFAULT),
                     f"H5Acreate2 {name}")
            _ok(L.H5Awrite(at, mtype, buf.ctypes.data), f"H5Awrite {name}")
            L.H5Aclose(at)
            L.H5Sclose(sp)
    finally:
        L.H5Fclose(f)


def read_h5(path):
    """(datasets, attrs) of a file written in this format: every grid and
    face dataset present and every attribute of write_grid."""
    L = lib()
    f = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen")
    out, attrs = {}, {}
    try:
        for name in PRIM_NAMES + CONS_NAMES[1:5] + FACE_NAMES:
            if name in out or L.H5Lexists(f, name.encode(), _H5P_DEFAULT) <= 0:
                continue
            ds = _ok(L.H5Dopen2(f, name.encode(), _H5P_DEFAULT), f"H5Dopen2 {name}")
            sp = L.H5Dget_space(ds)
            nd = L.H5Sget_simple_extent_ndims(sp)
            dims = (_hsize * nd)()
            L.H5Sget_simple_extent_dims(sp, dims, None)
            a = np.empty(tuple(dims[q] for q in range(nd)), dtype=np.float64)
            _ok(L.H5Dread(ds, _g("H5T_NATIVE_DOUBLE_g"), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                          a.ctypes.data), f"H5Dread {name}")
            L.H5Sclose(sp)
            L.H5Dclose(ds)
            out[name] = a
        for name in ("time", "step", "dt", "xf", "yf", "zf"):
            if L.H5Aexists(f, name.encode()) <= 0:
                continue
            at = L.H5Aopen(f, name.encode(), _H5P_DEFAULT)
            sp = L.H5Aget_space(at)
            nd = L.H5Sget_simple_extent_ndims(sp)
            dims = (_hsize * max(nd, 1))()
            if nd:
                L.H5Sget_simple_extent_dims(sp, dims, None)
            if name == "step":
                buf = np.empty((), dtype=np.int64)
                _ok(L.H5Aread(at, _g("H5T_NATIVE_LLONG_g"), buf.ctypes.data), "H5Aread")
                attrs[name] = int(buf)
            else:
                buf = np.empty(tuple(dims[q] for q in range(nd)), dtype=np.float64)
                _ok(L.H5Aread(at, _g("H5T_NATIVE_DOUBLE_g"), buf.ctypes.data), "H5Aread")
                attrs[name] = float(buf) if nd == 0 else buf
            L.H5Sclose(sp)
            L.H5Aclose(at)
    finally:
        L.H5Fclose(f)
    return out, attrs


def grid_datasets(shape, prim, bfld, cons=None, write_ghost=False):
    """Datasets of write_grid (output.pyx:286-361) from full-grid arrays:
    the primitive (or, given ``cons``, conserved) cell variables and the
    face fields, over the active cells or, with ghosts, the whole grid."""
    sel = (slice(None),) * 3 if write_ghost else shape.active()
    cell = prim if cons is None else cons
    names = PRIM_NAMES if cons is None else CONS_NAMES
    d = {names[n]: cell[n][sel] for n in range(9)}
    d.update({FACE_NAMES[n]: bfld[n][sel] for n in range(3)})
    return d


def face_coords(shape):
    """xf, yf, zf: the face coordinates lf of each axis, Ntot + 1 values
    (src/coordinates/coord.pyx:201-257; uniform faces start at 0 on the
    first active face, absent axes span [lmin, lmax])."""
    out = []
    for a in range(3):
        n = shape.ntot[a]
        if n == 1:
            out.append(np.array([shape.lmin[a], shape.lmax[a]], dtype=np.float64))
        else:
            dx = (shape.lmax[a] - shape.lmin[a]) / shape.nact[a]
            out.append((np.arange(n + 1) - shape.ng) * dx)
    return out
