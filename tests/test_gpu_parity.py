"""FP64 parity of the HIP path against the CPU oracle, through the C ABI.

Same seeded initial state on both sides, N steps, then every active cell of
prim and every active face of bfld is compared.  Tolerance: relative L1
<= 1e-12 per variable (the stated FP64 bar); both sides keep IEEE operation
order without contraction, so the differences are expected to be zero and
the tests report whether they are.
"""
import os
import sys

import numpy as np
import pytest

from bonzo_amd import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l1(a, b, scale=0.0):
    """sum|a - b| / sum|b|.  When scale is given and sum|b| < NOISE * scale,
    b is round-off noise of a quantity that vanishes in exact arithmetic (a
    transverse velocity or field component of a planar problem), and the
    difference is measured against scale instead of against that noise."""
    den = float(np.sum(np.abs(b)))
    if den < NOISE * scale:
        den = scale
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


# Below NOISE times the density's L1 norm a variable is round-off noise.
NOISE = 1e-9


def run_pair(prob, steps=None, tmax=None, rsolver="hlld", reconstr="linear", bc=(0,) * 6,
             dt_fms_ref=True, cour=0.4, precision="fp64", char_proj=False, hydro=False,
             ct_upwind=False, tintegr="vl"):
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    s, prim, bfld = prob
    order = 0 if reconstr == "const" else 1
    ng = 9 if tintegr == "rk3" else 2 if order == 0 else 3
    if s.ng != ng:  # rebuild the arrays with the grid's ghost layers
        prim, bfld = _regrid(s, prim, bfld, ng)
        s = P.GridShape(s.nx, s.ny, s.nz, ng, s.lmin, s.lmax)
    o = OracleSim(s.nx, s.ny, s.nz, reconstr_order=order, rsolver=rsolver, bc=bc, cour=cour,
                  dt_fms_ref=dt_fms_ref, char_proj=char_proj, hydro=hydro, ct_upwind=ct_upwind,
                  tintegr=tintegr)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    g = Context(s.nx, s.ny, s.nz, rsolver=rsolver, reconstr=reconstr, bc=bc, cour=cour,
                dt_fms_ref=dt_fms_ref, precision=precision, char_proj=char_proj, hydro=hydro,
                ct_upwind=ct_upwind, tintegr=tintegr)
    g.upload(prim, bfld)
    g.init_state()
    if steps is not None:
        o.advance_steps(steps)
        for _ in range(steps):
            g.run_steps(1)
    else:
        o.advance(tmax)
        g.advance(tmax)
    gp, gb, gc = g.download(cons=True)
    t, n, dt = g.time()
    res = dict(s=s, o=o, g=g, gp=gp, gb=gb, gc=gc, t=t, n=n, dt=dt)
    return res


def _regrid(s, prim, bfld, ng):
    s2 = P.GridShape(s.nx, s.ny, s.nz, ng, s.lmin, s.lmax)
    p2, b2 = s2.empty()
    a1 = s.active()
    a2 = s2.active()
    p2[(slice(None),) + a2] = prim[(slice(None),) + a1]
    b2[(slice(None),) + a2] = bfld[(slice(None),) + a1]
    return p2, b2


def compare(res, label):
    s, o = res["s"], res["o"]
    sk, sj, si = s.active()
    scale = float(np.sum(np.abs(o.prim[0][sk, sj, si])))
    worst = 0.0
    bitwise = True
    for v in range(9):
        a = res["gp"][v][sk, sj, si]
        b = o.prim[v][sk, sj, si]
        r = rel_l1(a, b, scale)
        bitwise &= bool(np.array_equal(a, b))
        worst = max(worst, r)
        assert r <= TOL, f"{label}: prim[{v}] rel L1 {r:.3e}"
    for c in range(3):
        a = res["gb"][c][sk, sj, si]
        b = o.bfld[c][sk, sj, si]
        r = rel_l1(a, b, scale)
        bitwise &= bool(np.array_equal(a, b))
        worst = max(worst, r)
        assert r <= TOL, f"{label}: bfld[{c}] rel L1 {r:.3e}"
    for v in range(9):
        a = res["gc"][v][sk, sj, si]
        b = o.cons[v][sk, sj, si]
        assert rel_l1(a, b, scale) <= TOL, f"{label}: cons[{v}]"
    say(f"{label}: worst rel L1 {worst:.3e}, bitwise={bitwise}")
    return bitwise


def test_pencil_solvers_match_oracle():
    from bonzo_amd._lib import load
    from oracle import oracle as orc
    from tests import mathref as M
    from tests.devmem import DeviceArray
    say("case: pencil hll/hlld/hllt, 4096 faces")
    rng = np.random.default_rng(21)
    wl, wr, bx = M.random_states(rng, 4096, bx_min=0.0)
    wl[1] += rng.uniform(-3, 3, 4096)
    wr[1] += rng.uniform(-3, 3, 4096)
    for rs, rid in (("hll", 0), ("hlld", 2), ("hllt", 4)):
        ref = orc.riemann(rs, wl, wr, bx, 5.0 / 3.0)
        bufs = [DeviceArray(a) for a in (wl, wr, bx, np.zeros((9, 4096)))]
        rc = load().bnz_gpu_flux_pencils(rid, *(b.ptr for b in bufs), 4096, 5.0 / 3.0)
        assert rc == 0
        got = bufs[3].get()
        for b in bufs:
            b.free()
        ref[6] = 0.0
        assert np.array_equal(got, ref), f"{rs}: max diff {np.max(np.abs(got - ref))}"


def test_rsolver_seam_matches_oracle_on_host_rows():
    """bnz_hll_flux / bnz_hlld_flux / bnz_hllt_flux have RSolverFunc's
    signature (host row pointers, faces i1..i2, integrator.pxd:76-79): on
    the same rows they write what the oracle's pencil solvers write, bit for
    bit, only at faces i1..i2, and a face with NaN states on both sides
    keeps its old flux row (the reference leaves an unmatched face
    untouched).  A NaN on one side only does not stop the solver: fmin/fmax
    of the wave-speed estimates drop the NaN, and the face takes the other
    side's region (fluxes_c.c:38-40)."""
    from bonzo_amd._lib import load
    from oracle import oracle as orc
    from oracle.oracle import _ptr, _rows
    from tests import mathref as M
    say("case: RSolverFunc seam hll/hlld/hllt on host rows, 2048 faces, i1=5, i2=2040")
    rng = np.random.default_rng(22)
    n, i1, i2 = 2048, 5, 2040
    wl, wr, bx = M.random_states(rng, n, bx_min=0.0)
    wl[1] += rng.uniform(-3, 3, n)
    wr[1] += rng.uniform(-3, 3, n)
    wl[:, 100] = np.nan  # NaN on both sides: no region matches, the row stays
    wr[:, 100] = np.nan
    wl[:, 101] = np.nan  # NaN on the left only: the speeds come from the right
    L = load()
    for rs in ("hll", "hlld", "hllt"):
        sentinel = 7.25
        want = np.full((9, n), sentinel)
        fn_o = {"hll": orc.lib().or_hll_flux, "hlld": orc.lib().or_hlld_flux,
                "hllt": orc.lib().or_hllt_flux}[rs]
        fn_o(_rows(want), _rows(wl), _rows(wr), _ptr(bx), i1, i2, 5.0 / 3.0)
        got = np.full((9, n), sentinel)
        getattr(L, f"bnz_{rs}_flux")(_rows(got), _rows(wl), _rows(wr), _ptr(bx), i1, i2,
                                     5.0 / 3.0)
        want[6, i1:i2 + 1] = got[6, i1:i2 + 1]  # normal-field slot: no consumer reads it
        assert np.all(got[:, :i1] == sentinel) and np.all(got[:, i2 + 1:] == sentinel)
        if rs != "hll":  # hll writes every face (NaN here), as the reference's does
            assert np.all(got[:, 100] == sentinel), rs
            assert np.all(np.isfinite(got[:, 101])), rs
        assert np.array_equal(got, want, equal_nan=True), \
            f"{rs}: max diff {np.nanmax(np.abs(got - want))}"


def test_hydro_hllc_seam_on_six_rows():
    """bnz_hllc_flux has the hydro build's RSolverFunc rows (rho, vx, vy, vz,
    p, scalar; fluxes_c.c:230-327 with NMODE = 6) and no bx (NULL): its six
    flux rows equal the oracle's HLLC on the same states with a zero field,
    bit for bit, only at faces i1..i2; nothing beyond row 5 is touched."""
    from bonzo_amd._lib import load
    from oracle import oracle as orc
    from oracle.oracle import _ptr, _rows
    from tests import mathref as M
    say("case: RSolverFunc seam hllc on 6 hydro rows, 1500 faces, i1=3, i2=1490")
    rng = np.random.default_rng(23)
    n, i1, i2 = 1500, 3, 1490
    wl, wr, _ = M.random_states(rng, n, bx_min=0.0)
    wl[1] += rng.uniform(-3, 3, n)
    wr[1] += rng.uniform(-3, 3, n)
    wl[6:] = 0.0
    wr[6:] = 0.0
    sentinel = 7.25
    want = np.full((9, n), sentinel)
    orc.lib().or_hllc_flux(_rows(want), _rows(wl), _rows(wr), _ptr(np.zeros(n)), i1, i2,
                           5.0 / 3.0)
    wl6 = np.ascontiguousarray(wl[:6])
    wr6 = np.ascontiguousarray(wr[:6])
    got = np.full((6, n), sentinel)
    load().bnz_hllc_flux(_rows(got), _rows(wl6), _rows(wr6), None, i1, i2, 5.0 / 3.0)
    assert np.all(got[:, :i1] == sentinel) and np.all(got[:, i2 + 1:] == sentinel)
    assert np.array_equal(got, want[:6]), f"max diff {np.max(np.abs(got - want[:6]))}"
    assert np.any(got[0, i1:i2 + 1] != 0.0)


@pytest.mark.parametrize("prec", ["fp64", "fp64fma", "fp32c"])
def test_2d_emf_placement_does_not_change_the_run(prec):
    """On 2D grids the edge EMFs are formed either in launches of their own
    or inside the update launches (emf_chunk 1 / 2; by default by grid
    size): the same expressions (in FP32C with the same in-wall zeros), so
    the same state bit for bit, and the IEEE unit equal to the oracle either
    way (periodic Orszag-Tang and a blast between reflecting walls)."""
    from bonzo_amd._lib import Context
    say(f"case: 2D EMF placement (own launches / in the updates), {prec}")
    for name, (s, prim, bfld), bc, steps in (
            ("ot48", P.orszag_tang(48), (0,) * 6, 12),
            ("blast40", P.blast2d(40, seed=11), (2, 2, 2, 2, 0, 0), 10)):
        out = []
        for mode in (1, 2):
            g = Context(s.nx, s.ny, 1, bc=bc, precision=prec)
            g.set_tuning(0, mode, 0)
            g.upload(prim, bfld)
            g.init_state()
            g.run_steps(steps)
            out.append(g.download() + (g.time(),))
            g.close()
        (pa, ba, ta), (pb, bb, tb) = out
        assert np.array_equal(pa, pb) and np.array_equal(ba, bb) and ta == tb, name
        if prec == "fp64":
            res = run_pair((s, prim, bfld), steps=steps, bc=bc)
            assert np.array_equal(res["gp"], pa) and np.array_equal(res["gb"], ba), name
            assert compare(res, name)


def test_orszag_tang_2d_bitwise():
    say("case: Orszag-Tang 64^2, 20 steps, hlld+plm, periodic")
    res = run_pair(P.orszag_tang(64), steps=20)
    assert compare(res, "OT64")


def test_orszag_tang_128_to_half_bitwise():
    """The reference's own CPU case: Orszag-Tang 128^2 run to t = 0.5 through
    advance(t_max) (about 400 steps, the last one clipped to t_max): state,
    time and step count equal the oracle's bit for bit."""
    say("case: Orszag-Tang 128^2 to t=0.5, hlld+plm, periodic")
    res = run_pair(P.orszag_tang(128), tmax=0.5)
    assert compare(res, "OT128 t=0.5")
    assert res["t"] == res["o"].time == 0.5
    assert res["n"] == res["o"].step


def test_turbulence_3d():
    say("case: turbulence 24^3, 6 steps, hlld+plm, periodic")
    res = run_pair(P.turbulence(24, seed_b=31, seed_v=32), steps=6)
    compare(res, "turb24")


def test_turbulence_3d_uniform_widths():
    """32 cells on [0, 1] give cell widths of exactly 2^-5 on every axis, so
    the update takes its geometry from kernel arguments (the uniform-grid
    kernels); the result must still be the oracle's bit for bit."""
    say("case: turbulence 32^3 (uniform widths), 5 steps, hlld+plm, periodic")
    res = run_pair(P.turbulence(32, seed_b=33, seed_v=34), steps=5)
    assert compare(res, "turb32")


def test_turbulence_256_headline_workload_bitwise():
    """The measured workload itself: 256^3 decaying MHD turbulence, seeded as
    the benchmark seeds it, two VL steps (PLM + HLLD + CT, periodic); every
    active cell and face of the FP64 path equals the oracle's bit for bit,
    and the FP32-compensated path is within 1e-5 relative L1 per variable
    with max |div B| <= 1e-12.  The oracle runs its OpenMP sweeps on 16 host
    threads (about 10 GB of host memory)."""
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    say("case: turbulence 256^3 (the benchmark's workload), 2 steps, hlld+plm, periodic, "
        "FP64 and FP32C")
    s, prim, bfld = P.turbulence(256)
    g = Context(256, 256, 256)
    g.upload(prim, bfld)
    g.init_state()
    g.run_steps(2)
    gp, gb = g.download()
    t, n, _ = g.time()
    g.close()
    h = Context(256, 256, 256, precision="fp32c")
    h.upload(prim, bfld)
    h.init_state()
    h.run_steps(2)
    hp, hb = h.download()
    divb = h.diag()["divb_max"]
    h.close()
    o = OracleSim(256, 256, 256, nthreads=16)
    o.prim[:] = prim
    o.bfld[:] = bfld
    del prim, bfld
    o.init()
    o.advance_steps(2)
    assert n == o.step == 2 and t == o.time
    sk, sj, si = s.active()
    for v in range(9):
        assert np.array_equal(gp[v][sk, sj, si], o.prim[v][sk, sj, si]), f"prim[{v}]"
    for c in range(3):
        assert np.array_equal(gb[c][sk, sj, si], o.bfld[c][sk, sj, si]), f"bfld[{c}]"
    worst = max([rel_l1(hp[v][sk, sj, si], o.prim[v][sk, sj, si]) for v in range(9)] +
                [rel_l1(hb[c][sk, sj, si], o.bfld[c][sk, sj, si]) for c in range(3)])
    say(f"  FP32C: worst rel L1 {worst:.3e}, max div B {divb:.3e}")
    assert worst <= 1e-5 and divb <= 1e-12
    o.close()


def test_brio_wu_1d_outflow_to_tmax():
    say("case: Brio-Wu 1D 400 cells, outflow, to t=0.08 with the tmax clip")
    res = run_pair(P.brio_wu(400, axis=0, transverse=0), tmax=0.08, bc=(1,) * 6)
    compare(res, "BW1D")
    assert res["t"] == res["o"].time
    assert res["n"] == res["o"].step


@pytest.mark.parametrize("axis", [1, 2])
def test_brio_wu_3d_outflow_along_axis(axis):
    say(f"case: Brio-Wu along axis {axis} in 3D, outflow, 8 steps")
    res = run_pair(P.brio_wu(32, axis=axis, transverse=4), steps=8, bc=(1,) * 6)
    compare(res, f"BW3D-{axis}")


def test_hll_solver_2d():
    say("case: OT 32^2 with hll, 12 steps")
    res = run_pair(P.orszag_tang(32), steps=12, rsolver="hll")
    compare(res, "OT32-hll")


def test_donor_cell_reconstruction_3d():
    say("case: turbulence 16^3 with reconstr=const (ng=2), 5 steps")
    res = run_pair(P.turbulence(16, seed_b=41, seed_v=42), steps=5, reconstr="const")
    compare(res, "turb16-const")


def test_dt_per_axis_fast_speed():
    say("case: OT 32^2 with per-axis fast speed in dt, 10 steps")
    res = run_pair(P.orszag_tang(32), steps=10, dt_fms_ref=False)
    compare(res, "OT32-dtfix")


def test_new_dt_matches_oracle():
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    say("case: new_dt on turbulence 20^3")
    s, prim, bfld = P.turbulence(20, seed_b=51, seed_v=52)
    o = OracleSim(20, 20, 20)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    g = Context(20, 20, 20)
    g.upload(prim, bfld)
    g.init_state()
    assert g.new_dt() == o.new_dt()


def test_divb_and_conservation_128():
    from bonzo_amd._lib import Context
    say("case: turbulence 128^3, 5 steps, div B and conservation via bnz_diag")
    s, prim, bfld = P.turbulence(128)
    g = Context(128, 128, 128)
    g.upload(prim, bfld)
    g.init_state()
    d0 = g.diag()
    g.run_steps(5)
    d1 = g.diag()
    say(f"  divB0={d0['divb_max']:.3e} divB5={d1['divb_max']:.3e} "
        f"dmass={d1['mass'] - d0['mass']:.3e} dE={d1['energy'] - d0['energy']:.3e}")
    assert d1["divb_max"] <= 1e-12 * max(1.0, d1["bmax"] * 128)
    assert abs(d1["mass"] - d0["mass"]) <= 1e-12 * abs(d0["mass"])
    assert abs(d1["energy"] - d0["energy"]) <= 1e-12 * abs(d0["energy"])
    for q in range(3):
        assert abs(d1["mom"][q] - d0["mom"][q]) <= 1e-13
    assert d1["nonfinite"] == 0


@pytest.mark.parametrize("chunks", [(0, 0, 0), (1, 1, 1), (7, 5, 3), (64, 64, 64), (0, 16, 4),
                                    (2, 2, 2), (26, 26, 26)])
def test_work_shapes_do_not_change_results(chunks):
    """Every work shape of the 3D flux, EMF and update kernels (flat, 64x4
    tiles, slab columns of any height, including partial and single columns)
    gives the oracle's result bit for bit."""
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    say(f"case: turbulence 20^3, shapes {chunks}, 4 steps")
    s, prim, bfld = P.turbulence(20, seed_b=61, seed_v=62)
    o = OracleSim(20, 20, 20)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    o.advance_steps(4)
    g = Context(20, 20, 20)
    g.set_tuning(*chunks)
    g.upload(prim, bfld)
    g.init_state()
    g.run_steps(4)
    gp, gb = g.download()
    sk, sj, si = s.active()
    assert np.array_equal(gp[:, sk, sj, si], o.prim[:, sk, sj, si])
    assert np.array_equal(gb[:, sk, sj, si], o.bfld[:, sk, sj, si])


@pytest.mark.parametrize("rs,order,cp", [("hlld", 1, False), ("hllt", 1, False),
                                         ("hll", 1, False), ("hlld", 0, False),
                                         ("hlld", 1, True)])
def test_whole_tiles_match_the_oracle(rs, order, cp):
    """Work shape 1 (64 x 4 tiles along z) on a box of whole and partial
    tiles (a turbulent state cut to 70 x 12 x 8, periodic): the IEEE build
    stays the oracle bit for bit for every solver, both orders, and with
    characteristic limiting."""
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    nx, ny, nz = 70, 12, 8
    ng = 3 if order else 2
    say(f"case: turbulence cut to {nx}x{ny}x{nz}, shapes (1, 1, 1), {rs} order {order} "
        f"char_proj {cp}, 3 steps")
    s, prim, bfld = P.turbulence(nx, ng=ng, seed_b=63, seed_v=64)
    sl = (slice(None), slice(0, nz + 2 * ng), slice(0, ny + 2 * ng), slice(0, nx + 2 * ng))
    prim, bfld = np.ascontiguousarray(prim[sl]), np.ascontiguousarray(bfld[sl])
    o = OracleSim(nx, ny, nz, reconstr_order=order, rsolver=rs, char_proj=cp)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    o.advance_steps(3)
    g = Context(nx, ny, nz, rsolver=rs, reconstr="linear" if order else "const", char_proj=cp)
    g.set_tuning(1, 1, 1)
    g.upload(prim, bfld)
    g.init_state()
    g.run_steps(3)
    gp, gb = g.download()
    g.close()
    act = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    assert np.array_equal(gp[act], o.prim[act])
    assert np.array_equal(gb[act], o.bfld[act])


@pytest.mark.parametrize("dims", [(70, 20, 6), (33, 9, 4)])
def test_tiles_fma(dims):
    """Work shape 1 on boxes of several tile columns and rows with partial
    ones: the IEEE build's tiled run is its flat run bit for bit, and the FMA
    build stays within its bar (1e-12 relative L1 per variable) of the IEEE
    run in both shapes (the compiler fuses each kernel's products on its own,
    so the FMA shapes may differ in the last bits)."""
    from bonzo_amd._lib import Context
    nx, ny, nz = dims
    say(f"case: turbulence cut to {nx}x{ny}x{nz}, fp64/fp64fma, shapes (0,0,0) vs (1,0,0), 6 steps")
    s, prim, bfld = P.turbulence(max(dims), seed_b=65, seed_v=66)
    sl = (slice(None), slice(0, nz + 6), slice(0, ny + 6), slice(0, nx + 6))
    prim, bfld = np.ascontiguousarray(prim[sl]), np.ascontiguousarray(bfld[sl])
    res = {}
    for prec in ("fp64", "fp64fma"):
        for shapes in ((0, 0, 0), (1, 0, 0)):
            g = Context(nx, ny, nz, precision=prec)
            g.set_tuning(*shapes)
            g.upload(prim, bfld)
            g.init_state()
            g.run_steps(6)
            res[prec, shapes] = g.download()[0][:, 3:-3, 3:-3, 3:-3]
            g.close()
    ref = res["fp64", (0, 0, 0)]
    assert np.array_equal(res["fp64", (1, 0, 0)], ref)
    for shapes in ((0, 0, 0), (1, 0, 0)):
        got = res["fp64fma", shapes]
        for v in range(9):
            den = max(np.abs(ref[v]).sum(), 1e-300)
            l1 = np.abs(got[v] - ref[v]).sum() / den
            assert l1 <= 1e-12, (shapes, v, l1)


def test_work_shapes_non_cubic():
    """Slab columns on a non-cubic grid with uneven column counts."""
    from bonzo_amd._lib import Context
    say("case: Brio-Wu along y on 6x40x6, outflow, shapes (5, 7, 3), 10 steps")
    res = {}
    for shapes in ((0, 0, 0), (5, 7, 3), (1, 1, 1)):
        s, prim, bfld = P.brio_wu(40, axis=1, transverse=6)
        g = Context(s.nx, s.ny, s.nz, bc=(1,) * 6)
        g.set_tuning(*shapes)
        g.upload(prim, bfld)
        g.init_state()
        g.run_steps(10)
        res[shapes] = g.download()
        g.close()
    for shapes in ((5, 7, 3), (1, 1, 1)):
        assert np.array_equal(res[shapes][0], res[(0, 0, 0)][0])
        assert np.array_equal(res[shapes][1], res[(0, 0, 0)][1])


def _blast_guide_field(n):
    """Blast with a uniform guide field along z only: no field crosses the
    walls of a closed box (a field normal to a reflecting wall is
    incompatible with it: the boundary zeroes it on the wall face)."""
    s, prim, bfld = P.blast2d(n, seed=21)
    bfld[0][:] = 0.0
    bfld[1][:] = 0.0
    bfld[2][s.active()] = 1.0
    return s, prim, bfld


def _tube_tangential_field(n):
    """Brio-Wu tube with the normal field removed (walls at both ends)."""
    s, prim, bfld = P.brio_wu(n, axis=0, transverse=0)
    bfld[0][:] = 0.0
    return s, prim, bfld


@pytest.mark.parametrize("case", ["blast2d_walls", "turb3d_mixed", "tube1d_walls", "wall2d_y"])
def test_reflecting_boundaries_bitwise(case):
    """Reflecting walls (mirror ghosts, sign-flipped normal velocity and
    field, zeroed wall face of the face-normal field) against the oracle,
    alone and mixed with periodic and outflow sides, on states whose field
    does not cross the walls."""
    from tests.test_oracle_pins import _wall_state
    if case == "blast2d_walls":
        say("case: blast 2D 48^2 with a z guide field in a reflecting box, 60 steps")
        res = run_pair(_blast_guide_field(48), steps=60, bc=(2,) * 6)
    elif case == "turb3d_mixed":
        say("case: turbulence 16^3, x periodic, y reflect, z outflow, 8 steps")
        res = run_pair(P.turbulence(16, seed_b=41, seed_v=42), steps=8, bc=(0, 0, 2, 2, 1, 1))
    elif case == "tube1d_walls":
        say("case: 1D tube 200 cells, tangential field, reflecting ends, 300 steps")
        res = run_pair(_tube_tangential_field(200), steps=300, bc=(2, 2, 0, 0, 0, 0))
    else:
        say("case: 2D state with walls in y (face-normal field zero there), 20 steps")
        res = run_pair(_wall_state(24, 20, 1), steps=20, bc=(0, 0, 2, 2, 0, 0))
    assert compare(res, case)
    sk, sj, si = res["s"].active()
    assert np.all(np.isfinite(res["gp"][:, sk, sj, si]))


def test_hllt_grid_bitwise():
    """HLLT through the whole step (5 steps: its energy flux inside the fan
    is undefined in the reference and set from zero star energies, which
    drives the pressure negative after a few more steps)."""
    say("case: Orszag-Tang 32^2, hllt + plm, 5 steps")
    res = run_pair(P.orszag_tang(32), steps=5, rsolver="hllt")
    assert compare(res, "OT32 hllt")


def test_history_row_matches_oracle(tmp_path):
    """bnz_history against the oracle's history row, and
    Simulation.write_history's file format.  The device sums in another
    order, so each column is held to 1e-12 of the scale its round-off has:
    the mean of |summand| (means of rho, energies, momenta, B), and for the
    rms columns sqrt(mean B^2 - mean(B)^2) the error that a round-off of
    mean B^2 leaves in the root, mean(B^2) / (2 rms)."""
    from bonzo_amd._lib import Context
    from oracle.oracle import OracleSim
    say("case: history row, turbulence 24^3 after 5 steps; hst.txt format")
    s, prim, bfld = P.turbulence(24, seed_b=51, seed_v=52)
    o = OracleSim(24, 24, 24)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    o.advance_steps(5)
    g = Context(24, 24, 24)
    g.upload(prim, bfld)
    g.init_state()
    g.run_steps(5)
    hg, ho = g.history(), o.history()
    g.close()
    assert hg[0] == ho[0] == 5 and hg[1] == ho[1]
    sk, sj, si = s.active()
    w = o.prim[:, sk, sj, si]
    rho, vel, bc = w[0], w[1:4], w[6:9]
    ek = 0.5 * rho * np.sum(vel ** 2, axis=0)
    em = 0.5 * np.sum(bc ** 2, axis=0)
    etot = ek + em + w[4] / (5.0 / 3.0 - 1.0)
    b2 = np.mean(bc ** 2, axis=(1, 2, 3))
    rms = ho[13:16]
    scale = np.concatenate([
        [0.0, 0.0, np.mean(np.abs(rho)), np.mean(ek), np.mean(etot), np.mean(em),
         np.mean(etot)],
        np.mean(np.abs(rho * vel), axis=(1, 2, 3)),
        np.mean(np.abs(bc), axis=(1, 2, 3)),
        b2 / (2.0 * rms)])
    assert np.all(rms > 0)
    assert np.all(np.abs(hg - ho) <= 1e-12 * scale), (hg - ho, scale)
    from bonzo_amd.simulation import Simulation
    cfg = {"computation": {"Nx": 16, "Ny": 16, "Nz": 1, "tmax": 0.01, "cour": 0.4,
                           "rsolver": "hlld", "reconstr": "linear"},
           "physics": {"gam": 5.0 / 3.0}, "problem": {"name": "otvortex"}}
    sim = Simulation(config=cfg)
    path = tmp_path / "hst.txt"
    sim.write_history(str(path))
    sim.advance(0.01)
    sim.write_history(str(path))
    sim.close()
    rows = [ln for ln in path.read_text().splitlines()]
    assert len(rows) == 2 and all(len(r) == 16 * 14 for r in rows)
    vals = [float(x) for x in rows[1].split()]
    assert len(vals) == 16 and vals[0] >= 1 and abs(vals[1] - 0.01) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_save_and_load_state_resumes_bitwise(prec):
    """A context that loads a saved state continues bit for bit like the one
    that saved it (including the dt the pending CFL maximum gives)."""
    from bonzo_amd._lib import Context
    say(f"case: state save/load, turbulence 20^3, {prec}, 6 + 5 steps")
    s, prim, bfld = P.turbulence(20, seed_b=91, seed_v=92)
    a = Context(20, 20, 20, precision=prec)
    a.upload(prim, bfld)
    a.init_state()
    a.run_steps(6)
    st = a.save_state()
    a.run_steps(5)
    b = Context(20, 20, 20, precision=prec)
    b.load_state(st)
    b.run_steps(5)
    assert a.time() == b.time()
    assert np.array_equal(a.save_state(), b.save_state())
    pa, ba = a.download()
    pb, bb = b.download()
    assert np.array_equal(pa, pb) and np.array_equal(ba, bb)
    other = Context(20, 20, 16, precision=prec)
    with pytest.raises(Exception):
        other.load_state(st)
    for c in (a, b, other):
        c.close()


def test_simulation_restart_file(tmp_path):
    from bonzo_amd.simulation import Simulation
    say("case: Simulation.save_restart / load_restart, OT 32^2")
    cfg = {"computation": {"Nx": 32, "Ny": 32, "Nz": 1, "tmax": 0.1, "cour": 0.4,
                           "rsolver": "hlld", "reconstr": "linear"},
           "physics": {"gam": 5.0 / 3.0}, "problem": {"name": "otvortex"}}
    a = Simulation(config=cfg)
    a.advance(0.02)
    path = str(tmp_path / "rst.npz")
    a.save_restart(path)
    a.advance(0.05)
    b = Simulation(config=cfg)
    b.load_restart(path)
    assert abs(b.integr.time - 0.02) < 1e-15
    b.advance(0.05)
    assert np.array_equal(a.prim, b.prim) and np.array_equal(a.bfld, b.bfld)
    with np.load(path) as f:
        assert f["rho"].shape == a.prim[0].shape
        assert abs(float(f["time"]) - 0.02) < 1e-15 and int(f["step"]) > 0
        assert set(Simulation.GRID_NAMES + Simulation.FACE_NAMES) <= set(f.files)
    a.close()
    b.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_simulation_hdf5_restart_resumes_bitwise(tmp_path, dims):
    """The reference's own restart file (rst/rst_grid.h5: primitive
    variables and face fields of the active cells, time and step;
    src/io/restart.pyx:23-86): a run restarted from it continues bit for
    bit like the uninterrupted run, and the grid output file holds the
    state of the run (src/io/output.pyx:248-385)."""
    from bonzo_amd import h5io
    from bonzo_amd.simulation import Simulation
    say(f"case: Simulation HDF5 restart, {dims}D")
    n = 32 if dims == 2 else 16
    cfg = {"computation": {"Nx": n, "Ny": n, "Nz": n if dims == 3 else 1, "tmax": 0.1,
                           "cour": 0.4, "rsolver": "hlld", "reconstr": "linear"},
           "physics": {"gam": 5.0 / 3.0},
           "problem": {"name": "otvortex" if dims == 2 else "turbulence"}}
    a = Simulation(usr_dir=str(tmp_path), config=cfg)
    a.advance(0.02)
    rst = a.write_restart_h5()
    assert rst == os.path.join(str(tmp_path), "rst", "rst_grid.h5")
    out = a.write_grid(0.01)
    assert os.path.basename(out) == "grid_002.h5"
    ds, attrs = h5io.read_h5(out)
    act = a.grid.shape.active()
    assert np.array_equal(ds["p"], a.prim[4][act]) and np.array_equal(ds["bzf"], a.bfld[2][act])
    assert attrs["time"] == a.integr.time and attrs["step"] == a.integr.step
    a.advance(0.05)
    b = Simulation(usr_dir=str(tmp_path), config=cfg)
    b.read_restart_h5()
    assert b.integr.time == attrs["time"] and b.integr.step == attrs["step"]
    b.advance(0.05)
    assert a.integr.step == b.integr.step
    assert np.array_equal(a.prim[:, act[0], act[1], act[2]], b.prim[:, act[0], act[1], act[2]])
    assert np.array_equal(a.bfld[:, act[0], act[1], act[2]], b.bfld[:, act[0], act[1], act[2]])
    a.close()
    b.close()


def test_advance_chunks_and_idle_steps_change_nothing():
    """advance() issues steps in chunks and reads the time while the next
    chunk runs; steps that start at t_max get dt = 0 and must leave the state,
    time and step count alone.  Interrupted runs (max_steps) and repeated
    calls at t_max give the uninterrupted run bit for bit, in both
    precisions."""
    from bonzo_amd._lib import Context
    say("case: Brio-Wu 64, OT 32^2, turbulence 16^3 to t=0.03: max_steps pieces, repeat at "
        "t_max, resume")
    for prec in ("fp64", "fp32c"):
        for s, prim, bfld in (P.brio_wu(64, transverse=0), P.orszag_tang(32), P.turbulence(16)):
            ref = Context(s.nx, s.ny, s.nz, precision=prec)
            ref.upload(prim, bfld)
            ref.init_state()
            n_ref = ref.advance(0.03)
            st_ref, t_ref = ref.save_state(), ref.time()
            g = Context(s.nx, s.ny, s.nz, precision=prec)
            g.upload(prim, bfld)
            g.init_state()
            n = g.advance(0.03, max_steps=3) + g.advance(0.03, max_steps=5) + g.advance(0.03)
            assert n == n_ref and g.time()[:2] == t_ref[:2]
            assert g.time()[0] == 0.03
            assert g.advance(0.03) == 0                 # already there: nothing happens
            assert np.array_equal(g.save_state(), st_ref)
            n2 = g.advance(0.04)
            assert n2 > 0 and g.time()[0] == 0.04
            ref.close()
            g.close()


def test_upload_waits_for_queued_steps_and_needs_init():
    """upload() is ordered after the steps still queued on the context's
    stream and leaves a state that must be initialised before the next step;
    after init_state the context runs exactly like a fresh one, in both
    precisions."""
    from bonzo_amd._lib import BNZ_ESTATE, BnzError, Context
    say("case: re-upload after 12 queued steps, turbulence 20^3 (two seeds), fp64 and fp32c")
    s, prim0, bfld0 = P.turbulence(20, seed_b=31, seed_v=32)
    _, prim1, bfld1 = P.turbulence(20, seed_b=41, seed_v=42)
    for prec in ("fp64", "fp32c"):
        a = Context(20, 20, 20, precision=prec)
        a.upload(prim0, bfld0)
        a.init_state()
        a.run_steps(12)                  # stream-ordered, no host synchronisation
        a.upload(prim1, bfld1)
        with pytest.raises(BnzError) as e:
            a.run_steps(1)
        assert e.value.code == BNZ_ESTATE
        a.init_state()
        a.run_steps(4)
        b = Context(20, 20, 20, precision=prec)
        b.upload(prim1, bfld1)
        b.init_state()
        b.run_steps(4)
        assert a.time() == b.time()
        assert np.array_equal(a.save_state(), b.save_state())
        a.close()
        b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_diagnostics_and_retuning_do_not_change_the_run(prec):
    """diag(), history() and set_tuning() between steps leave the run bit for
    bit as it is: the next dt still comes from the CFL maximum the last
    corrector reduced (in FP32C from the FP64 state, which a fresh reduction
    over the FP32 prim would not reproduce)."""
    from bonzo_amd._lib import Context
    say(f"case: diag/history/set_tuning between 8 steps, turbulence 20^3, {prec}")
    s, prim, bfld = P.turbulence(20, seed_b=51, seed_v=52)
    a = Context(20, 20, 20, precision=prec)
    a.upload(prim, bfld)
    a.init_state()
    a.run_steps(8)
    b = Context(20, 20, 20, precision=prec)
    b.upload(prim, bfld)
    b.init_state()
    shapes = ((0, 0, 0), (4, 0, 8), (1, 1, 1), (64, 64, 64))
    for q in range(8):
        b.run_steps(1)
        d = b.diag()
        assert d["nonfinite"] == 0 and d["divb_max"] < 1e-12
        b.history()
        b.set_tuning(*shapes[q % 4])
    assert a.time() == b.time()
    assert np.array_equal(a.save_state(), b.save_state())
    a.close()
    b.close()


FMA_CASES = {
    "ot128_to_half": dict(prob=lambda: P.orszag_tang(128), tmax=0.5),
    "turb64_100": dict(prob=lambda: P.turbulence(64), steps=100),
    "turb24_hll_const": dict(prob=lambda: P.turbulence(24, seed_b=3, seed_v=4), steps=20,
                             rsolver="hll", reconstr="const"),
    "brio_wu_1d": dict(prob=lambda: P.brio_wu(400, axis=0, transverse=0), tmax=0.08,
                       bc=(1,) * 6),
    # HLLT's zero fan-region energy flux loses positivity after 6 steps here
    "hllt_ot32": dict(prob=lambda: P.orszag_tang(32), steps=5, rsolver="hllt"),
    # the guide-field blast: a seeded blast whose field crosses the walls
    # loses positivity by t = 0.02 in the reference's scheme (no floor)
    "blast2d_walls": dict(prob=lambda: _blast_guide_field(48), tmax=0.05, bc=(2,) * 6),
}


@pytest.mark.parametrize("case", sorted(FMA_CASES))
def test_fp64_fma_mode_within_1e12_of_oracle(case):
    """The FP64 step with contracted multiply-adds (BNZ_PREC_FP64_FMA): the
    same algorithm with each a*b+c rounded once, so it is not bit for bit the
    oracle; every variable stays within the FP64 bar of 1e-12 relative L1,
    with the same step count and time on the runs to t_max."""
    kw = dict(FMA_CASES[case])
    prob = kw.pop("prob")()
    say(f"case: FP64-FMA mode vs oracle, {case}")
    res = run_pair(prob, precision="fp64fma", **kw)
    o, (sk, sj, si) = res["o"], res["s"].active()
    # a case is only meaningful while the oracle's own run stays physical
    assert np.all(np.isfinite(o.prim[:, sk, sj, si]))
    assert o.prim[0][sk, sj, si].min() > 0 and o.prim[4][sk, sj, si].min() > 0
    if "tmax" in kw:
        assert res["n"] == res["o"].step and res["t"] == res["o"].time
    compare(res, f"fp64fma {case}")
    res["o"].close()
    res["g"].close()


CHAR_CASES = {
    "ot64_2d": dict(prob=lambda: P.orszag_tang(64), steps=25),
    "turb24_3d": dict(prob=lambda: P.turbulence(24, seed_b=3, seed_v=4), steps=6),
    "brio_wu_1d": dict(prob=lambda: P.brio_wu(400, axis=0, transverse=0), tmax=0.08,
                       bc=(1,) * 6),
    "brio_wu_y_3d": dict(prob=lambda: P.brio_wu(64, axis=1, transverse=4), steps=20,
                         bc=(0, 0, 1, 1, 0, 0)),
    "hll_blast_walls": dict(prob=lambda: _blast_guide_field(40), steps=40, bc=(2,) * 6,
                            rsolver="hll"),
}


@pytest.mark.parametrize("precision", ["fp64", "fp64fma"])
@pytest.mark.parametrize("case", sorted(CHAR_CASES))
def test_characteristic_limiting_matches_oracle(case, precision):
    """charact_proj=1: PLM slopes limited in characteristic variables
    (src/coordinates/reconstr_c.c:64-80 with the projections of
    src/mhd/eos_c.h:100-280), on the 3D, 2D and fused 1D kernels: the FP64
    path equals the oracle (1e-12 relative L1, expected bitwise), and the FMA
    mode stays within 1e-12."""
    kw = dict(CHAR_CASES[case])
    prob = kw.pop("prob")()
    say(f"case: charact_proj=1, {case}, {precision}")
    res = run_pair(prob, precision=precision, char_proj=True, **kw)
    o, (sk, sj, si) = res["o"], res["s"].active()
    assert np.all(np.isfinite(o.prim[:, sk, sj, si]))
    assert o.prim[0][sk, sj, si].min() > 0 and o.prim[4][sk, sj, si].min() > 0
    bitwise = compare(res, f"char {case} {precision}")
    if precision == "fp64":
        assert bitwise
    res["o"].close()
    res["g"].close()


HYDRO_CASES = {
    "sod_1d_hllc": dict(prob=lambda: P.sod(400), tmax=0.15, bc=(1,) * 6, rsolver="hllc"),
    "sod_1d_hll": dict(prob=lambda: P.sod(400), tmax=0.15, bc=(1,) * 6, rsolver="hll"),
    "sod_y_3d_hllc": dict(prob=lambda: P.sod(64, axis=1, transverse=4), steps=30,
                          bc=(0, 0, 1, 1, 0, 0), rsolver="hllc"),
    "blast_2d_hllc": dict(prob=lambda: P.hydro_blast(64), steps=40, rsolver="hllc"),
    "blast_2d_hllc_const": dict(prob=lambda: P.hydro_blast(48), steps=30, rsolver="hllc",
                                reconstr="const"),
    "blast_3d_hllc_walls": dict(prob=lambda: P.hydro_blast(24, dims=3), steps=12, rsolver="hllc",
                                bc=(2,) * 6),
    "blast_3d_hll": dict(prob=lambda: P.hydro_blast(24, dims=3, seed=8), steps=12,
                         rsolver="hll"),
}


@pytest.mark.parametrize("precision", ["fp64", "fp64fma"])
@pytest.mark.parametrize("case", sorted(HYDRO_CASES))
def test_hydro_build_matches_oracle(case, precision):
    """The reference's hydro build (MFIELD=0; HLLC of fluxes_c.c:230-327 or
    HLL with the sound speed, hydro dt): the FP64 path equals the oracle's
    hydro mode bit for bit, the FMA mode within 1e-12, on 1D (fused step),
    2D and 3D grids with outflow, periodic and reflecting sides; the field
    stays exactly zero."""
    kw = dict(HYDRO_CASES[case])
    prob = kw.pop("prob")()
    say(f"case: hydro build, {case}, {precision}")
    res = run_pair(prob, precision=precision, hydro=True, **kw)
    o, (sk, sj, si) = res["o"], res["s"].active()
    assert np.all(np.isfinite(o.prim[:, sk, sj, si]))
    assert o.prim[0][sk, sj, si].min() > 0 and o.prim[4][sk, sj, si].min() > 0
    if "tmax" in kw:
        assert res["n"] == o.step and res["t"] == o.time
    bitwise = compare(res, f"hydro {case} {precision}")
    if precision == "fp64" and not bitwise:
        # The passive scalar of the blast decays to tiny values ahead of its
        # front; where two of its differences multiply to less than 2^-970
        # (differences below ~1e-146) the limiter divides such a numerator:
        # the one range in which the device's Newton division departs from
        # the IEEE quotient, by an ulp (by design, bnz_phys.hpp).  Values
        # of that size may differ; everything else is bitwise.
        a = np.concatenate([res["gp"][:, sk, sj, si].ravel(), res["gb"][:, sk, sj, si].ravel()])
        b = np.concatenate([o.prim[:, sk, sj, si].ravel(), o.bfld[:, sk, sj, si].ravel()])
        off = a != b
        assert np.all(np.maximum(np.abs(a[off]), np.abs(b[off])) < 1e-140), (
            np.max(np.abs(a[off])), np.max(np.abs(b[off])))
    assert np.all(res["gb"] == 0.0) and np.all(res["gp"][6:9] == 0.0)
    o.close()
    res["g"].close()


def test_hydro_pencil_hllc_and_zero_field_upload():
    """bnz_gpu_flux_pencils with BNZ_RS_HLLC equals the oracle's HLLC bit
    for bit; a hydro context refuses a face field that is not zero."""
    from bonzo_amd._lib import BnzError, Context, load
    from oracle import oracle as orc
    from tests import mathref as M
    from tests.devmem import DeviceArray
    say("case: pencil hllc, 4096 faces; hydro upload with a non-zero face field")
    rng = np.random.default_rng(61)
    wl, wr, bx = M.random_hydro_states(rng, 4096)
    wl[1] += rng.uniform(-3, 3, 4096)
    ref = orc.riemann("hllc", wl, wr, bx, 5.0 / 3.0)
    bufs = [DeviceArray(a) for a in (wl, wr, bx, np.zeros((9, 4096)))]
    assert load().bnz_gpu_flux_pencils(1, *(b.ptr for b in bufs), 4096, 5.0 / 3.0) == 0
    got = bufs[3].get()
    for b in bufs:
        b.free()
    ref[6] = 0.0
    assert np.array_equal(got, ref)
    s, prim, bfld = P.sod(64)
    g = Context(s.nx, s.ny, s.nz, rsolver="hllc", bc=(1,) * 6, hydro=True)
    bfld[1][s.active()] = 1e-3
    with pytest.raises(BnzError):
        g.upload(prim, bfld)
    bfld[:] = 0.0
    g.upload(prim, bfld)
    g.init_state()
    g.run_steps(3)
    _, gb = g.download()
    assert np.all(gb == 0.0)
    g.close()


UPWIND_CASES = {
    "turb24_3d": dict(prob=lambda: P.turbulence(24, seed_b=5, seed_v=6), steps=8),
    "turb20_3d_const_hll": dict(prob=lambda: P.turbulence(20, seed_b=7, seed_v=8), steps=8,
                                rsolver="hll", reconstr="const"),
    "ot64_2d": dict(prob=lambda: P.orszag_tang(64), steps=30),
    "brio_wu_z_3d": dict(prob=lambda: P.brio_wu(64, axis=2, transverse=4), steps=20,
                         bc=(0, 0, 0, 0, 1, 1)),
    "blast_walls_2d": dict(prob=lambda: _blast_guide_field(40), steps=30, bc=(2,) * 6),
    "ot32_to_tmax": dict(prob=lambda: P.orszag_tang(32), tmax=0.1),
}


@pytest.mark.parametrize("precision", ["fp64", "fp64fma"])
@pytest.mark.parametrize("case", sorted(UPWIND_CASES))
def test_upwind_corner_emf_matches_oracle(case, precision):
    """ct_upwind = 1: the upwind corner EMF (interp_ee2, ct.pyx:229-358 with
    its Ex indices corrected) on 3D (flat and tiled work shapes by the
    default tuning) and 2D grids, periodic, outflow and reflecting: the IEEE
    unit equals the oracle bit for bit, the FMA unit within 1e-12, and div B
    stays at round-off."""
    kw = dict(UPWIND_CASES[case])
    prob = kw.pop("prob")()
    say(f"case: upwind corner EMF, {case}, {precision}")
    res = run_pair(prob, precision=precision, ct_upwind=True, **kw)
    o, (sk, sj, si) = res["o"], res["s"].active()
    assert np.all(np.isfinite(o.prim[:, sk, sj, si]))
    assert o.prim[0][sk, sj, si].min() > 0 and o.prim[4][sk, sj, si].min() > 0
    if "tmax" in kw:
        assert res["n"] == o.step and res["t"] == o.time
    bitwise = compare(res, f"upwind {case} {precision}")
    if precision == "fp64":
        assert bitwise
    d = res["g"].diag()
    assert d["divb_max"] <= 1e-11
    o.close()
    res["g"].close()


RK3_CASES = {
    "turb24_3d": dict(prob=lambda: P.turbulence(24, seed_b=9, seed_v=10), steps=6),
    "turb20_3d_const_hll": dict(prob=lambda: P.turbulence(20, seed_b=11, seed_v=12), steps=6,
                                rsolver="hll", reconstr="const"),
    "ot64_to_tmax": dict(prob=lambda: P.orszag_tang(64), tmax=0.05),
    "brio_wu_1d": dict(prob=lambda: P.brio_wu(256, axis=0, transverse=0), tmax=0.05,
                       bc=(1,) * 6),
    "blast_walls_2d": dict(prob=lambda: _blast_guide_field(40), steps=15, bc=(2,) * 6),
    "sod_hydro_hllc": dict(prob=lambda: P.sod(200), tmax=0.1, bc=(1,) * 6, rsolver="hllc",
                           hydro=True),
    "turb16_3d_upwind": dict(prob=lambda: P.turbulence(16, seed_b=13, seed_v=14), steps=5,
                             ct_upwind=True),
}


@pytest.mark.parametrize("precision", ["fp64", "fp64fma"])
@pytest.mark.parametrize("case", sorted(RK3_CASES))
def test_rk3_matches_oracle(case, precision):
    """tintegr = rk3: the three-stage SSP Runge-Kutta of _rk3.pyx (ng = 9, no
    ghost fill between the stages) on 1D, 2D and 3D grids, with PLM and
    donor cell, HLLD, HLL and the hydro build's HLLC, the upwind corner EMF,
    periodic, outflow and reflecting sides: the IEEE unit equals the oracle
    bit for bit (time and step count on runs to t_max), the FMA unit within
    1e-12."""
    kw = dict(RK3_CASES[case])
    prob = kw.pop("prob")()
    say(f"case: rk3, {case}, {precision}")
    res = run_pair(prob, precision=precision, tintegr="rk3", **kw)
    o, (sk, sj, si) = res["o"], res["s"].active()
    assert res["s"].ng == 9
    assert np.all(np.isfinite(o.prim[:, sk, sj, si]))
    assert o.prim[0][sk, sj, si].min() > 0 and o.prim[4][sk, sj, si].min() > 0
    if "tmax" in kw:
        assert res["n"] == o.step and res["t"] == o.time
    bitwise = compare(res, f"rk3 {case} {precision}")
    if precision == "fp64":
        assert bitwise
    o.close()
    res["g"].close()


@pytest.mark.parametrize("prec", ["fp64", "fp64fma", "fp32c"])
def test_advance_graph_equals_eager_steps(prec):
    """advance() replays a captured chunk of eight steps once a full chunk
    has started and ended on the same buffers, CFL partial count and work
    shapes (step graph on); with the step graph off it issues every launch
    itself.  Both give the same state, time and step count bit for bit, on
    1D (fused step), 2D and 3D grids and with RK3, including a run that ends
    inside a chunk (max_steps), one that reaches t_max, one resumed from a
    reloaded state (fresh CFL partials) and one after a change of work
    shapes."""
    from bonzo_amd._lib import Context
    say(f"case: advance graph vs eager launches, {prec}")
    probs = [("tube", P.brio_wu(200, axis=0, transverse=0), (1,) * 6, "vl"),
             ("ot", P.orszag_tang(48), (0,) * 6, "vl"),
             ("turb", P.turbulence(16, seed_b=3, seed_v=4), (0,) * 6, "vl")]
    if prec != "fp32c":
        probs.append(("turb_rk3", P.turbulence(16, seed_b=5, seed_v=6, ng=9), (0,) * 6, "rk3"))
    for name, (s, prim, bfld), bc, tint in probs:
        out = []
        for on in (True, False):
            g = Context(s.nx, s.ny, s.nz, bc=bc, precision=prec, tintegr=tint)
            g.set_step_graph(on)
            g.upload(prim, bfld)
            g.init_state()
            g.advance(0.03, max_steps=21)
            g.advance(0.045)
            g.load_state(g.save_state())
            g.advance(0.06)
            if s.nz > 1:
                g.set_tuning(8, 0, 16)
            g.advance(0.075)
            out.append(g.download() + (g.time(),))
            g.close()
        (pa, ba, ta), (pb, bb, tb) = out
        assert ta == tb, (name, ta, tb)
        assert np.array_equal(pa, pb) and np.array_equal(ba, bb), name
