"""User-boundary specs that restate the built-in flags, for tests: a flag-3
side given the copies of periodic (flag 0) or outflow (flag 1) must fill its
ghosts exactly as the flag does (src/bc/grid_bc_funcs.pyx:36-196)."""
import numpy as np

from bonzo_amd.user_bc import Mapped


def builtin_as_user(ntot, ng, flag, axis, side):
    """Mapped spec of side (axis, side) that copies as flag 0 (periodic: the
    active layers at the far end) or 1 (outflow: the first or last active
    layer) does."""
    nact = ntot[axis] - 2 * ng

    def source(i, j, k):
        idx = [np.array(i), np.array(j), np.array(k)]
        x = idx[axis]
        if flag == 0:
            idx[axis] = x + nact if side == 0 else x - nact
        else:
            idx[axis] = np.full_like(x, ng if side == 0 else ntot[axis] - ng - 1)
        return tuple(idx)

    return Mapped(source)


def arrays(spec, ntot, ng, dims, axis, side):
    from bonzo_amd.user_bc import side_arrays
    return side_arrays(spec, ntot, ng, dims, axis, side)
