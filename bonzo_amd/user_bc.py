"""User boundaries (flag 3) as data.

The reference leaves a flag-3 side to code in the problem file: a function
per side, assigned in set_bc_grid_ptrs_user (src/problem/problem.pxd:13,
called from the GridBC hook at src/bc/grid_bc.pyx:266-268), which writes
the side's ghost cells each time the boundaries are applied.  The one
problem that uses it, st2d.pyx:198-372, shows both kinds of fill such code
does: fixed inflow states (x1, x2) and copies from chosen cells (y1, y2:
the opposite rows shifted along x).

Here a side is described once, as data, and the device applies it on every
fill (bnz_set_user_bc):

  Fixed(prim, faces, rows)   fixed values on the cells rows() selects
  Mapped(source, rows)       each selected cell copies every variable of the
                             cell source(i, j, k) names

The slab of a side is its ng ghost layers along the axis times the full
extent (ghosts included) of the other axes, x fastest; cells no spec
selects keep their values.  Sides are applied in the reference's order z1,
z2, y1, y2, x1, x2, each reading what the earlier ones wrote.
"""

import numpy as np

NFIX = 12  # rho, vx, vy, vz, p, scalar, bx, by, bz (cell) + the x, y, z faces


def slab(ntot, ng, axis, side):
    """(i, j, k) index arrays of a side's slab, in slab order (x fastest).
    ntot is (Ntot_x, Ntot_y, Ntot_z)."""
    lo = [0, 0, 0]
    n = list(ntot)
    lo[axis] = ntot[axis] - ng if side else 0
    n[axis] = ng
    k, j, i = np.meshgrid(np.arange(lo[2], lo[2] + n[2]), np.arange(lo[1], lo[1] + n[1]),
                          np.arange(lo[0], lo[0] + n[0]), indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def _active_rows(ntot, ng, dims, axis):
    """rows() selecting the cells active along every axis but the side's own:
    the extent the reference's st2d x-side fills cover (st2d.pyx:222-224)."""
    def sel(i, j, k):
        m = np.ones(i.shape, dtype=bool)
        for a, x in enumerate((i, j, k)):
            if a == axis or a >= dims:
                continue
            m &= (x >= ng) & (x < ntot[a] - ng)
        return m
    return sel


class Fixed:
    """Fixed values: prim = (rho, vx, vy, vz, p, scalar, bx, by, bz) at cell
    centres, faces = (bx, by, bz) on the cell's lower faces (default: the
    cell-centred field).  rows(i, j, k) -> bool mask of the cells to set
    (default: all of the slab; "active" for the cells active along the other
    axes)."""

    def __init__(self, prim, faces=None, rows=None):
        self.prim = np.asarray(prim, dtype=np.float64)
        if self.prim.shape != (9,):
            raise ValueError("Fixed: prim takes 9 values (rho, vx, vy, vz, p, scalar, bx, by, bz)")
        self.faces = self.prim[6:9] if faces is None else np.asarray(faces, dtype=np.float64)
        if self.faces.shape != (3,):
            raise ValueError("Fixed: faces takes 3 values")
        self.rows = rows

    def entries(self, ntot, ng, dims, axis, side):
        i, j, k = slab(ntot, ng, axis, side)
        rows = _active_rows(ntot, ng, dims, axis) if self.rows == "active" else self.rows
        m = np.ones(i.shape, dtype=bool) if rows is None else np.asarray(rows(i, j, k), bool)
        src = np.where(m, -1, -2).astype(np.int32)
        fixed = np.empty((i.size, NFIX))
        fixed[:, :9] = self.prim
        fixed[:, 9:] = self.faces
        return src, fixed


class Mapped:
    """Copies: source(i, j, k) -> (i2, j2, k2), index arrays of the cell each
    slab cell takes every variable from.  rows as for Fixed."""

    def __init__(self, source, rows=None):
        self.source = source
        self.rows = rows

    def entries(self, ntot, ng, dims, axis, side):
        i, j, k = slab(ntot, ng, axis, side)
        i2, j2, k2 = (np.broadcast_to(np.asarray(x, dtype=np.int64), i.shape)
                      for x in self.source(i, j, k))
        for a, x in enumerate((i2, j2, k2)):
            if x.min(initial=0) < 0 or x.max(initial=0) >= ntot[a]:
                raise ValueError(f"Mapped: source index out of the grid on axis {a}")
        src = ((k2 * ntot[1] + j2) * ntot[0] + i2).astype(np.int32)
        rows = _active_rows(ntot, ng, dims, axis) if self.rows == "active" else self.rows
        if rows is not None:
            src = np.where(np.asarray(rows(i, j, k), bool), src, -2).astype(np.int32)
        return src, None


def side_arrays(spec, ntot, ng, dims, axis, side):
    """(src int32, fixed float64 [n][12] or None) of a spec for one side; a
    (src, fixed) pair passes through."""
    if hasattr(spec, "entries"):
        return spec.entries(ntot, ng, dims, axis, side)
    src, fixed = spec
    return (np.ascontiguousarray(src, dtype=np.int32),
            None if fixed is None else np.ascontiguousarray(fixed, dtype=np.float64))
