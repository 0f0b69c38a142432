"""Batched boxes against single grids, through the C ABI.

A box of a batch runs the single-grid arithmetic with its own dt, time and
t_max, so every box must equal, bit for bit, the single-grid run of the same
initial state to the same t_max (same step count, same final time); the
single-grid path is tied to the oracle by the FP64 parity tests.  Cases mix
box sizes, dimensionalities, boundary flags and t_max, run with and without
the captured hipGraph, and stop early with max_steps.
"""
import sys

import numpy as np
import pytest

from bonzo_amd import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def box_desc(s, bc, tmax):
    return dict(nx=s.nx, ny=s.ny, nz=s.nz, bc=bc, lmin=s.lmin, lmax=s.lmax, tmax=tmax)


def single(s, prim, bfld, bc, tmax, rsolver, reconstr, max_steps=0, precision="fp64"):
    from bonzo_amd._lib import Context
    g = Context(s.nx, s.ny, s.nz, rsolver=rsolver, reconstr=reconstr, bc=bc,
                precision=precision)
    g.upload(prim, bfld)
    g.init_state()
    n = g.advance(tmax, max_steps)
    p_, b_ = g.download()
    t = g.time()[0]
    g.close()
    return p_, b_, t, n


def run_batch(cases, rsolver="hlld", reconstr="linear", use_graph=True, max_steps=0,
              precision="fp64"):
    """cases: list of ((shape, prim, bfld), bc, tmax)."""
    from bonzo_amd._lib import Batch
    b = Batch([box_desc(c[0][0], c[1], c[2]) for c in cases], rsolver=rsolver,
              reconstr=reconstr, precision=precision)
    for q, ((s, prim, bfld), _, _) in enumerate(cases):
        b.upload(q, prim, bfld)
    b.init_state()
    n = b.advance(max_steps=max_steps, use_graph=use_graph)
    t, steps = b.times()
    out = [b.download(q) for q in range(len(cases))]
    b.close()
    return out, t, steps, n


def check_against_single(cases, rsolver="hlld", reconstr="linear", use_graph=True,
                         precision="fp64"):
    out, t, steps, n = run_batch(cases, rsolver, reconstr, use_graph, precision=precision)
    say(f"  batch: {n} step launches, per-box steps {list(steps)}")
    for q, ((s, prim, bfld), bc, tmax) in enumerate(cases):
        p1, b1, t1, n1 = single(s, prim, bfld, bc, tmax, rsolver, reconstr,
                                precision=precision)
        assert steps[q] == n1, (q, steps[q], n1)
        assert t[q] == t1, (q, t[q], t1)
        assert np.array_equal(out[q][0], p1), (q, np.max(np.abs(out[q][0] - p1)))
        assert np.array_equal(out[q][1], b1), (q, np.max(np.abs(out[q][1] - b1)))
    return steps


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_tubes_bitwise_vs_single(use_graph):
    say(f"case: 7 Riemann tubes 64..700 cells, outflow, HLLD+PLM, graph={use_graph}")
    rng = np.random.Generator(np.random.PCG64(4004))
    lengths = [64, 700, 129, 333, 64, 517, 256]
    tmaxes = [0.1, 0.1, 0.05, 0.1, 0.02, 0.08, 0.1]
    cases = [(P.riemann_tube(n, rng), (1,) * 6, tm) for n, tm in zip(lengths, tmaxes)]
    steps = check_against_single(cases, use_graph=use_graph)
    assert len(set(steps)) > 2  # boxes finish at different steps


def test_batch_blast_2d_bitwise_vs_single():
    say("case: 3 periodic 2D blast boxes 48/72/40, HLLD+PLM+CT, uneven t_max")
    cases = [(P.blast2d(48, seed=11), (0,) * 6, 0.02),
             (P.blast2d(72, seed=12), (0,) * 6, 0.015),
             (P.blast2d(40, seed=13), (0,) * 6, 0.03)]
    check_against_single(cases)


def test_batch_mixed_dims_hll_const_bitwise():
    say("case: 1D + 2D + 3D boxes in one batch, HLL + donor cell, mixed flags")
    rng = np.random.Generator(np.random.PCG64(77))
    cases = [(P.riemann_tube(100, rng, ng=2), (1,) * 6, 0.05),
             (P.orszag_tang(40, ng=2), (0,) * 6, 0.05),
             (P.turbulence(16, ng=2, seed_b=5, seed_v=6), (0, 0, 1, 1, 0, 0), 0.05),
             (P.brio_wu(90, axis=1, ng=2), (1,) * 6, 0.04)]
    check_against_single(cases, rsolver="hll", reconstr="const")


def test_batch_mixed_dims_hlld_plm_bitwise():
    say("case: 3D turbulence 20^3 + OT 48^2 + tube 200 in one batch, HLLD+PLM")
    rng = np.random.Generator(np.random.PCG64(78))
    cases = [(P.turbulence(20, seed_b=7, seed_v=8), (0,) * 6, 0.03),
             (P.orszag_tang(48), (0,) * 6, 0.04),
             (P.riemann_tube(200, rng), (1,) * 6, 0.05)]
    check_against_single(cases)


def test_batch_max_steps_and_resume():
    """Stopping early with max_steps and resuming gives the same state as one
    uninterrupted run; graph and eager steps interleave."""
    say("case: 4 tubes, max_steps 5 eager, then graph to the end, vs one run")
    from bonzo_amd._lib import Batch
    rng = np.random.Generator(np.random.PCG64(99))
    cases = [(P.riemann_tube(n, rng), (1,) * 6, 0.05) for n in (80, 160, 96, 250)]
    ref, tref, sref, _ = run_batch(cases, use_graph=False)
    b = Batch([box_desc(c[0][0], c[1], c[2]) for c in cases])
    for q, ((s, prim, bfld), _, _) in enumerate(cases):
        b.upload(q, prim, bfld)
    b.init_state()
    assert b.advance(max_steps=5, use_graph=False) == 5
    t, steps = b.times()
    assert list(steps) == [5] * 4
    b.advance(max_steps=3, use_graph=True)   # odd eager count before: capture at parity 1
    b.advance(max_steps=1, use_graph=False)  # parity changes: graph is re-captured
    b.advance(use_graph=True)
    t, steps = b.times()
    assert np.array_equal(t, tref) and np.array_equal(steps, sref)
    for q in range(len(cases)):
        p_, b_ = b.download(q)
        assert np.array_equal(p_, ref[q][0]) and np.array_equal(b_, ref[q][1])
    b.close()


@pytest.mark.parametrize("precision,mix", [("fp64", "2d"), ("fp32c", "2d"), ("fp64", "1d"),
                                           ("fp32c", "1d")])
def test_batch_finished_boxes_across_pieces(precision, mix):
    """Boxes that reach t_max long before the others sit out many steps (the
    face buffers, and in an all-1D FP64 batch the prim buffers, swap without
    them); advancing in pieces of odd and even length, graph and eager, and
    reading a finished box in between, must give each box its uninterrupted
    state."""
    say(f"case: boxes finishing at different steps, pieces, {precision}, {mix}")
    from bonzo_amd._lib import Batch
    rng = np.random.Generator(np.random.PCG64(1234))
    if mix == "2d":
        cases = [(P.riemann_tube(64, rng), (1,) * 6, 0.01),
                 (P.orszag_tang(32), (0,) * 6, 0.005),
                 (P.riemann_tube(300, rng), (1,) * 6, 0.06),
                 (P.orszag_tang(48), (0,) * 6, 0.03)]
    else:
        cases = [(P.riemann_tube(64, rng), (1,) * 6, 0.01),
                 (P.riemann_tube(700, rng), (0,) * 6, 0.03),
                 (P.riemann_tube(300, rng), (1,) * 6, 0.06),
                 (P.riemann_tube(96, rng), (1,) * 6, 0.002)]
    ref, tref, sref, _ = run_batch(cases, use_graph=False, precision=precision)
    say(f"  reference steps {list(sref)}")
    b = Batch([box_desc(c[0][0], c[1], c[2]) for c in cases], precision=precision)
    for q, ((s, prim, bfld), _, _) in enumerate(cases):
        b.upload(q, prim, bfld)
    b.init_state()
    pieces = [(3, False), (5, True), (2, False), (7, False), (4, True)]
    checked = set()
    for it in range(400):
        n, g = pieces[it % len(pieces)]
        b.advance(max_steps=n, use_graph=g)
        t, steps = b.times()
        for q in range(len(cases)):
            if q not in checked and t[q] == tref[q]:
                p_, b_ = b.download(q)
                assert np.array_equal(p_, ref[q][0]) and np.array_equal(b_, ref[q][1]), q
                checked.add(q)
        if all(t[q] == tref[q] for q in range(len(cases))):
            break
    assert np.array_equal(steps, sref)
    for q in range(len(cases)):
        p_, b_ = b.download(q)
        assert np.array_equal(p_, ref[q][0]), q
        assert np.array_equal(b_, ref[q][1]), q
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_batch_launch_sizes_do_not_change_results(precision):
    """The workgroup caps of the stage launches (bnz_batch_set_tuning) only
    change how the work blocks are dealt out: one workgroup for everything,
    odd caps and the defaults give the same states bit for bit."""
    say(f"case: launch-size caps (1,1,1), (7,3,5), default; 1D + 2D + 3D boxes, {precision}")
    from bonzo_amd._lib import Batch, BnzError
    rng = np.random.Generator(np.random.PCG64(4242))
    cases = [(P.riemann_tube(300, rng), (1,) * 6, 0.03),
             (P.orszag_tang(40), (0,) * 6, 0.02),
             (P.turbulence(12, seed_b=21, seed_v=22), (0,) * 6, 0.01)]
    ref, tref, sref, _ = run_batch(cases, precision=precision)
    for caps in ((1, 1, 1), (7, 3, 5)):
        b = Batch([box_desc(c[0][0], c[1], c[2]) for c in cases], precision=precision)
        b.set_tuning(*caps)
        for q, ((s, prim, bfld), _, _) in enumerate(cases):
            b.upload(q, prim, bfld)
        b.init_state()
        b.advance()
        t, steps = b.times()
        assert np.array_equal(t, tref) and np.array_equal(steps, sref), caps
        for q in range(len(cases)):
            p_, b_ = b.download(q)
            assert np.array_equal(p_, ref[q][0]) and np.array_equal(b_, ref[q][1]), (caps, q)
        with pytest.raises(BnzError):
            b.set_tuning(-1, 0, 0)
        b.close()


def test_tube_sweep_conserves_and_reaches_tmax():
    """A slice of the 1D sweep workload (64 tubes of 64..4096 cells) runs to
    t=0.1 with the graph; every box ends exactly at t_max with finite state."""
    say("case: tube sweep 64 boxes, HLLD+PLM to t=0.1, graph")
    sweep = P.tube_sweep(64, seed=4004)
    cases = [(c, (1,) * 6, 0.1) for c in sweep]
    out, t, steps, n = run_batch(cases)
    say(f"  {n} step launches, steps per box {steps.min()}..{steps.max()}")
    assert np.all(t == 0.1)
    for q, ((s, _, _), _, _) in enumerate(cases):
        sk, sj, si = s.active()
        assert np.all(np.isfinite(out[q][0][:, sk, sj, si]))
        assert np.all(out[q][0][0][sk, sj, si] > 0) and np.all(out[q][0][4][sk, sj, si] > 0)


def test_batch_reflecting_boxes_bitwise_vs_single():
    say("case: reflecting blast 40^2, walled 2D state, tube with a wall, mixed flags")
    from tests.test_oracle_pins import _wall_state
    rng = np.random.Generator(np.random.PCG64(123))
    from tests.test_gpu_parity import _blast_guide_field
    tube = P.riemann_tube(150, rng)
    tube[2][0][:] = 0.0  # no normal field through the wall
    cases = [(_blast_guide_field(40), (2,) * 6, 0.02),
             (_wall_state(24, 20, 0), (2, 2, 0, 0, 0, 0), 0.03),
             (tube, (2, 1, 0, 0, 0, 0), 0.05),
             (P.turbulence(12, seed_b=3, seed_v=4), (1, 1, 2, 2, 0, 0), 0.02)]
    check_against_single(cases)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_fp32c_bitwise_vs_single_fp32c(use_graph):
    """FP32-compensated batches run the single-grid FP32C bodies: every box
    equals its FP32C single-grid run (tubes, 2D blasts, a 3D box, walls)."""
    say(f"case: FP32C batch of tubes, 2D blasts, 3D turbulence, walls; graph={use_graph}")
    from tests.test_gpu_parity import _blast_guide_field
    from tests.test_oracle_pins import _wall_state
    rng = np.random.Generator(np.random.PCG64(4005))
    cases = [(P.riemann_tube(n, rng), (1,) * 6, tm)
             for n, tm in ((96, 0.05), (300, 0.1), (64, 0.02))]
    walled = P.riemann_tube(150, rng)
    walled[2][0][:] = 0.0  # no normal field through the wall
    cases += [(walled, (2, 1, 0, 0, 0, 0), 0.05),
              (P.riemann_tube(128, rng), (0,) * 6, 0.03)]
    cases += [(P.blast2d(40, seed=14), (0,) * 6, 0.02),
              (_blast_guide_field(32), (2,) * 6, 0.03),
              (_wall_state(24, 20, 0), (2, 2, 0, 0, 0, 0), 0.03),
              (P.turbulence(12, seed_b=15, seed_v=16), (0,) * 6, 0.02)]
    check_against_single(cases, use_graph=use_graph, precision="fp32c")


@pytest.mark.parametrize("rsolver,reconstr,precision",
                         [("hlld", "linear", "fp64"), ("hllt", "linear", "fp64"),
                          ("hll", "const", "fp64"), ("hlld", "const", "fp64"),
                          ("hlld", "linear", "fp32c"), ("hll", "const", "fp32c")])
def test_batch_all_1d_fused_steps_bitwise_vs_single(rsolver, reconstr, precision):
    """A batch of 1D boxes only runs each step as one fused pass (predictor
    and corrector in LDS, state buffers swapped like the faces): every box
    equals its single-grid run, with outflow, periodic and walled tubes of
    lengths that do and do not fill the last work block."""
    say(f"case: all-1D batch, fused steps, {rsolver} + {reconstr}, {precision}")
    ng = 2 if reconstr == "const" else 3
    rng = np.random.Generator(np.random.PCG64(4008))
    walled = P.riemann_tube(251, rng, ng=ng)
    walled[2][0][:] = 0.0
    cases = [(P.riemann_tube(n, rng, ng=ng), bc, tm)
             for n, bc, tm in ((64, (1,) * 6, 0.05), (502, (0,) * 6, 0.02),
                               (245, (1,) * 6, 0.04), (1000, (1,) * 6, 0.01))]
    cases.append((walled, (2, 2, 0, 0, 0, 0), 0.03))
    check_against_single(cases, rsolver=rsolver, reconstr=reconstr, precision=precision)
    check_against_single(cases, rsolver=rsolver, reconstr=reconstr, use_graph=False,
                         precision=precision)


def test_batch_hllt_bitwise_vs_single():
    say("case: HLLT batch, PLM, tubes (outflow, periodic, walled) + a 2D box")
    rng = np.random.Generator(np.random.PCG64(4007))
    walled = P.riemann_tube(140, rng)
    walled[2][0][:] = 0.0
    cases = [(P.riemann_tube(200, rng), (1,) * 6, 0.05),
             (P.riemann_tube(96, rng), (0,) * 6, 0.03),
             (walled, (1, 2, 0, 0, 0, 0), 0.04),
             (P.orszag_tang(32), (0,) * 6, 0.02)]
    check_against_single(cases, rsolver="hllt")


def test_batch_fp32c_hll_const_bitwise_vs_single_fp32c():
    say("case: FP32C batch, HLL + donor cell (ng=2), 1D + 2D + 3D boxes, mixed flags")
    rng = np.random.Generator(np.random.PCG64(4006))
    cases = [(P.riemann_tube(120, rng, ng=2), (1,) * 6, 0.05),
             (P.orszag_tang(36, ng=2), (0,) * 6, 0.04),
             (P.turbulence(14, ng=2, seed_b=17, seed_v=18), (0, 0, 1, 1, 0, 0), 0.03)]
    check_against_single(cases, rsolver="hll", reconstr="const", precision="fp32c")
