"""HDF5 grid and restart files (bonzo_amd/h5io.py) in the reference's layout
(src/io/output.pyx:248-385, src/io/restart.pyx:23-86), checked through the
HDF5 library's own introspection: dataset names, shapes, type, chunking,
attributes, and a write/read round trip."""
import ctypes
import os

import numpy as np
import pytest

from bonzo_amd import h5io
from bonzo_amd import problems as P


def test_grid_file_layout_and_round_trip(tmp_path):
    s = P.GridShape(12, 10, 8)
    rng = np.random.default_rng(5)
    prim = rng.standard_normal((9,) + s.ntot[::-1])
    bfld = rng.standard_normal((3,) + s.ntot[::-1])
    ds = h5io.grid_datasets(s, prim, bfld)
    xf, yf, zf = h5io.face_coords(s)
    path = str(tmp_path / "out" / "grid_003.h5")
    h5io.write_h5(path, ds, dict(time=0.31, step=57, dt=1.5e-3, xf=xf, yf=yf, zf=zf))
    with open(path, "rb") as f:
        assert f.read(8) == b"\x89HDF\r\n\x1a\n"
    got, attrs = h5io.read_h5(path)
    names = list(h5io.PRIM_NAMES) + list(h5io.FACE_NAMES)
    assert sorted(got) == sorted(names)
    act = s.active()
    for n, name in enumerate(h5io.PRIM_NAMES):
        assert got[name].shape == (8, 10, 12)
        assert np.array_equal(got[name], prim[n][act])
    for n, name in enumerate(h5io.FACE_NAMES):
        assert np.array_equal(got[name], bfld[n][act])
    assert attrs["time"] == 0.31 and attrs["step"] == 57 and attrs["dt"] == 1.5e-3
    assert np.array_equal(attrs["xf"], xf) and len(attrs["xf"]) == s.ntot[0] + 1
    assert attrs["xf"][s.ng] == 0.0 and attrs["xf"][s.ng + 12] == pytest.approx(1.0)
    # the library's view: IEEE float64 little-endian, chunked (layout 2)
    L = h5io.lib()
    f = L.H5Fopen(path.encode(), 0, 0)
    d = L.H5Dopen2(f, b"rho", 0)
    t = L.H5Dget_type(d)
    assert L.H5Tequal(t, h5io._g("H5T_IEEE_F64LE_g")) > 0
    pl = L.H5Dget_create_plist(d)
    assert L.H5Pget_layout(pl) == 2
    L.H5Pclose(pl)
    L.H5Tclose(t)
    L.H5Dclose(d)
    a = L.H5Aopen(f, b"step", 0)
    t = L.H5Aget_type(a)
    assert L.H5Tequal(t, h5io._g("H5T_STD_I64LE_g")) > 0
    L.H5Tclose(t)
    L.H5Aclose(a)
    L.H5Fclose(f)


def test_conserved_names_and_ghosts(tmp_path):
    s = P.GridShape(6, 5, 1)
    prim = np.arange(9 * 11 * 12, dtype=float).reshape((9, 1, 11, 12))
    cons = -prim
    bfld = np.ones((3, 1, 11, 12))
    ds = h5io.grid_datasets(s, prim, bfld, cons=cons, write_ghost=True)
    path = str(tmp_path / "g.h5")
    h5io.write_h5(path, ds, dict(time=0.0, step=0, dt=0.0), chunked=False)
    got, _ = h5io.read_h5(path)
    for name in ("mx", "my", "mz", "en"):
        assert name in got and got[name].shape == (1, 11, 12)
    assert np.array_equal(got["en"], cons[4])
