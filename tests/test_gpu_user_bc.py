"""User boundaries (flag 3) on the device against the oracle, through the C
ABI: the oblique shock tube of src/problem/st2d.pyx with its four user
sides, user sides that restate the built-in flags, call errors, and the
Simulation path with the problem's own user boundaries."""
import sys

import numpy as np
import pytest

from bonzo_amd import problems as P
from tests import userbc_util as U

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l1(a, b):
    den = float(np.sum(np.abs(b)))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


def device_run(prob, bc, specs, steps, precision="fp64", graph=False, tmax=None):
    from bonzo_amd._lib import Context
    s, prim, bfld = prob[:3]
    g = Context(s.nx, s.ny, s.nz, bc=bc, lmax=s.lmax, precision=precision)
    for (a, d), spec in specs.items():
        g.set_user_bc(a, d, spec)
    g.set_step_graph(graph)
    g.upload(prim, bfld)
    g.init_state()
    if tmax is None:
        g.run_steps(steps)
    else:
        g.advance(tmax)
    p_, b_ = g.download()
    t = g.time()
    div = g.diag()["divb_max"]
    g.close()
    return p_, b_, t, div


def oracle_run(prob, bc, specs, steps):
    from oracle.oracle import OracleSim
    s, prim, bfld = prob[:3]
    o = OracleSim(s.nx, s.ny, s.nz, bc=tuple(bc), lmax=s.lmax)
    for (a, d), spec in specs.items():
        src, fixed = U.arrays(spec, o.ntot, o.ng, s.dims, a, d)
        o.set_user_bc(a, d, src, fixed)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    o.advance_steps(steps)
    return o


@pytest.mark.parametrize("prec", ["fp64", "fp64fma"])
def test_oblique_tube_user_sides_match_oracle(prec):
    """st2d's four user sides (fixed x states over the active rows, y rows
    copied with a 2 ny shift, clamped to the active columns), 30 steps on
    the 8:1 box: the IEEE path bit for bit with the oracle, the FMA path
    within 1e-12 relative L1; the fixed and copied ghost cells are part of
    the comparison.  div B stays at round-off of the corner potential's
    differences (the fixed x faces agree with the field while no wave has
    reached the x sides, which the 8:1 box ensures for 30 steps)."""
    say(f"case: oblique tube 128x16 user sides, 30 steps, {prec}")
    prob = P.oblique_tube(16, aspect=8)
    s, prim, bfld, user = prob
    bc = (3, 3, 3, 3, 0, 0)
    o = oracle_run(prob, bc, user, 30)
    gp, gb, t, div = device_run(prob, bc, user, 30, precision=prec)
    if prec == "fp64":
        assert np.array_equal(gp, o.prim) and np.array_equal(gb, o.bfld)
        assert t[0] == o.time and t[1] == 30
    else:
        sk, sj, si = s.active()
        for v in (0, 1, 2, 3, 4, 6, 7, 8):
            r = rel_l1(gp[v][sk, sj, si], o.prim[v][sk, sj, si])
            assert r <= 1e-12, (v, r)
        assert abs(t[0] - o.time) <= 1e-13 * o.time
    assert div <= 1e-11
    sk, sj, si = s.active()
    assert np.max(np.abs(o.prim[0][sk, sj, si] - prim[0][sk, sj, si])) > 1e-1
    o.close()


def test_oblique_tube_user_sides_rk3():
    """The same user sides with the RK3 integrator (ng = 9, so the slabs are
    nine layers deep and the y copies reach nine rows in): the device bit
    for bit with the oracle after 12 steps."""
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    say("case: oblique tube 128x16 user sides, RK3, 12 steps")
    s0, prim0, bfld0, _ = P.oblique_tube(16, aspect=8)
    s, prim, bfld, user = P.oblique_tube(16, ng=9, aspect=8)
    bc = (3, 3, 3, 3, 0, 0)
    o = OracleSim(s.nx, s.ny, 1, bc=bc, lmax=s.lmax, tintegr="rk3")
    g = Context(s.nx, s.ny, 1, bc=bc, lmax=s.lmax, tintegr="rk3")
    assert o.ng == g.ng == 9
    for (a, d), spec in user.items():
        src, fixed = U.arrays(spec, o.ntot, o.ng, s.dims, a, d)
        o.set_user_bc(a, d, src, fixed)
        g.set_user_bc(a, d, spec)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    g.upload(prim, bfld)
    g.init_state()
    dt = 0.5 * o.new_dt()
    for _ in range(12):
        o.rk3_step(dt)
    for _ in range(12):
        g.vl_step(dt)
    gp, gb = g.download()
    assert np.array_equal(gp, o.prim) and np.array_equal(gb, o.bfld)
    sk, sj, si = s.active()
    assert np.max(np.abs(o.prim[0][sk, sj, si] - prim[0][sk, sj, si])) > 1e-2
    g.close()
    o.close()


def test_oblique_tube_user_sides_fp32c():
    """FP32C with user sides (fixed rows converted to its conserved hi words
    and FP32 prim once): within the FP32C shock-tube bar of the FP64 path
    and the translation symmetry (i, j) -> (i - 2, j + 1) kept to FP32 level
    on the 8:1 box.  Two of st2d's sides overwrite an active face on every
    fill: x2 the last column's upper x face (a fixed value, :252-303) and y2
    the top row's upper y face (a copy of a bottom face 2 ny columns on,
    :332-357, from the last active column where that image is off the
    grid).  Both agree with the CT-evolved faces only as far as the
    far-field states do, and those carry FP32 rounding in FP32C, so div B
    there is at FP32 level; in every other active cell it is at round-off."""
    say("case: oblique tube 128x16 user sides, 30 steps, fp32c vs fp64")
    prob = P.oblique_tube(16, aspect=8)
    s, prim, bfld, user = prob
    bc = (3, 3, 3, 3, 0, 0)
    p64, b64, t64, _ = device_run(prob, bc, user, 30)
    p32, b32, t32, div = device_run(prob, bc, user, 30, precision="fp32c")
    sk, sj, si = s.active()
    e = [rel_l1(p32[v][sk, sj, si], p64[v][sk, sj, si]) for v in (0, 1, 2, 4, 6, 7)]
    say("  rel L1 " + " ".join(f"{x:.2e}" for x in e))
    assert max(e) <= 1e-4
    assert div <= 1e-5
    g = s.ng
    b = b32[:, 0]
    dv = ((b[0][g:g + s.ny, g + 1:g + s.nx + 1] - b[0][g:g + s.ny, g:g + s.nx]) +
          (b[1][g + 1:g + s.ny + 1, g:g + s.nx] - b[1][g:g + s.ny, g:g + s.nx])) / s.dx(0)
    say(f"  div B: top row {np.abs(dv[-1]).max():.2e}, last column "
        f"{np.abs(dv[:, -1]).max():.2e}, elsewhere {np.abs(dv[:-1, :-1]).max():.2e}")
    assert np.abs(dv[:-1, :-1]).max() <= 1e-11
    a = p32[0, 0, g:g + s.ny, g:g + s.nx]
    assert np.max(np.abs(a[:-1, 2:] - a[1:, :-2])) <= 1e-5


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("dims,flags", [(2, (1, 1, 0, 0, 0, 0)), (3, (1, 1, 0, 0, 1, 1)),
                                        (3, (0, 0, 0, 0, 0, 0))])
def test_user_sides_restating_builtin_flags_on_device(prec, dims, flags):
    """User sides given the copies of the built-in flags, side by side in
    the reference's order, against the built-in composed fill on the device:
    bit for bit where the restated sides are outflow; where an axis was
    periodic its ghost widths differ in the last bits (only periodic axes
    re-space them), so the runs agree to round-off (FP64) or to FP32 level
    (FP32C)."""
    say(f"case: user sides restating flags {flags}, {dims}D, {prec}")
    prob = P.orszag_tang(24) if dims == 2 else P.turbulence(12, seed_b=21, seed_v=22)
    s = prob[0]
    ref = device_run(prob, flags, {}, 8, precision=prec)
    nt = (s.nx + 2 * s.ng, s.ny + 2 * s.ng, s.nz + 2 * s.ng if dims == 3 else 1)
    bc = [3] * (2 * dims) + list(flags[2 * dims:])
    specs = {(q // 2, q % 2): U.builtin_as_user(nt, s.ng, flags[q], q // 2, q % 2)
             for q in range(2 * dims)}
    got = device_run(prob, bc, specs, 8, precision=prec)
    if 0 not in flags[:2 * dims]:
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert got[2] == ref[2]
    else:
        sk, sj, si = s.active()
        tol = 1e-13 if prec == "fp64" else 1e-6
        for v in range(9):
            r = rel_l1(got[0][v][sk, sj, si], ref[0][v][sk, sj, si])
            assert r <= tol or np.sum(np.abs(ref[0][v][sk, sj, si])) == 0, (v, r)


def test_user_bc_call_errors_and_reset():
    """Steps and fills need every flag-3 side set (BNZ_ESTATE); a side whose
    flag is not 3, an axis beyond the grid, a source off the grid or inside
    the side's own written cells, and -1 entries without fixed values are
    refused; so are boundaries that leave a ghost cell with a value read
    from a ghost cell a later side fills (the reference's st2d clamp into
    the x sides' ghost columns; x sides over the active rows only next to a
    periodic y axis, whose corner copies read the x ghosts); a refused side
    leaves the boundaries as they were; setting a side again takes effect
    at the next fill."""
    from bonzo_amd._lib import BNZ_EINVAL, BNZ_ESTATE, BnzError, Context
    from bonzo_amd.user_bc import Fixed, Mapped
    say("case: user boundary call errors, 32x8")
    prob = P.oblique_tube(8)
    s, prim, bfld, user4 = prob
    # x sides over the whole slab (with a periodic y axis the corners are theirs)
    user = {(0, 0): Fixed(user4[(0, 0)].prim), (0, 1): Fixed(user4[(0, 1)].prim)}
    st = Context(s.nx, s.ny, 1, bc=(3, 3, 3, 3, 0, 0), lmax=s.lmax)
    ntx = st.ntot[0]
    st.set_user_bc(0, 0, user4[(0, 0)])
    st.set_user_bc(0, 1, user4[(0, 1)])
    st.set_user_bc(1, 0, user4[(1, 0)])
    with pytest.raises(BnzError) as e:
        st.set_user_bc(1, 1, Mapped(lambda i, j, k: (np.minimum(i + 16, ntx - 1), j - 8, k)))
    assert e.value.code == BNZ_EINVAL and "later side" in str(e.value)
    with pytest.raises(BnzError) as e:
        st.init_state()
    assert e.value.code == BNZ_ESTATE
    st.set_user_bc(1, 1, user4[(1, 1)])
    st.upload(prim, bfld)
    st.init_state()
    st.close()
    pr = Context(s.nx, s.ny, 1, bc=(3, 3, 0, 0, 0, 0), lmax=s.lmax)
    pr.set_user_bc(0, 0, user4[(0, 0)])
    with pytest.raises(BnzError) as e:
        pr.set_user_bc(0, 1, user4[(0, 1)])
    assert e.value.code == BNZ_EINVAL and "later side" in str(e.value)
    pr.close()
    g = Context(s.nx, s.ny, 1, bc=(3, 3, 0, 0, 0, 0), lmax=s.lmax)
    g.upload(prim, bfld)
    with pytest.raises(BnzError) as e:
        g.init_state()
    assert e.value.code == BNZ_ESTATE and "x1" in str(e.value)
    g.set_user_bc(0, 0, user[(0, 0)])
    with pytest.raises(BnzError) as e:
        g.fill_ghosts()
    assert e.value.code == BNZ_ESTATE and "x2" in str(e.value)
    for axis, side, spec in ((1, 0, user4[(1, 0)]), (2, 0, user[(0, 0)]),
                             (0, 1, Mapped(lambda i, j, k: (i - 1, j, k)))):
        with pytest.raises(BnzError) as e:
            g.set_user_bc(axis, side, spec)
        assert e.value.code == BNZ_EINVAL
    n = 3 * g.ntot[1]
    for src, fixed in ((np.full(n, g.ntot[0] * g.ntot[1], np.int32), None),
                       (np.full(n, -1, np.int32), None)):
        with pytest.raises(BnzError) as e:
            g.set_user_bc(0, 1, (src, fixed))
        assert e.value.code == BNZ_EINVAL
    g.set_user_bc(0, 1, user[(0, 1)])
    g.init_state()
    g.run_steps(5)
    a = g.download()[0]
    # the same run with the right side's state changed after the init
    h = Context(s.nx, s.ny, 1, bc=(3, 3, 0, 0, 0, 0), lmax=s.lmax)
    h.set_user_bc(0, 0, user[(0, 0)])
    h.set_user_bc(0, 1, user[(0, 1)])
    h.upload(prim, bfld)
    h.init_state()
    w = user[(0, 1)].prim.copy()
    w[4] *= 2.0
    h.set_user_bc(0, 1, Fixed(w))
    h.run_steps(5)
    b = h.download()[0]
    g0 = s.ng
    assert np.array_equal(a[:, 0, g0:-g0, :s.nx // 2], b[:, 0, g0:-g0, :s.nx // 2])
    assert np.max(np.abs(a[4, 0, g0:-g0, -g0 - 1] - b[4, 0, g0:-g0, -g0 - 1])) > 1e-2
    g.close()
    h.close()


def test_simulation_st2d_user_boundaries():
    """Simulation with [physics] bc 3 on four sides and the st2d problem
    takes the problem's user boundaries (set_bc_grid_ptrs_user) and runs as
    the Context does, with and without step graphs."""
    from bonzo_amd.simulation import Simulation
    say("case: Simulation st2d 64x16 to t=0.05")
    cfg = {"computation": {"Nx": 64, "Ny": 16, "Nz": 1, "tmax": 0.05, "cour": 0.4,
                           "rsolver": "hlld", "reconstr": "linear", "xmax": 4.0, "ymax": 1.0},
           "physics": {"gam": 5.0 / 3.0, "bc_x1": 3, "bc_x2": 3, "bc_y1": 3, "bc_y2": 3},
           "problem": {"name": "st2d"}}
    sim = Simulation(config=cfg)
    n = sim.advance(0.05)
    prob = P.oblique_tube(16)
    for graph in (False, True):
        gp, gb, t, _ = device_run(prob, (3, 3, 3, 3, 0, 0), prob[3], None, graph=graph,
                                  tmax=0.05)
        assert np.array_equal(sim.prim, gp) and np.array_equal(sim.bfld, gb), graph
        assert t[1] == n
    sim.close()
