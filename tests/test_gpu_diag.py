"""Failure detection of bnz_diag, through the C ABI (data tests: states
that hold a non-finite cell and non-positive pressure or density are
uploaded and inspected; no step runs on them)."""
import sys

import numpy as np
import pytest

from bonzo_amd import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_diag_locates_the_first_failing_cell(prec):
    from bonzo_amd._lib import Context
    say(f"case: diag of an Orszag-Tang 32^2 state with a NaN, a p<0 and a rho<0 cell, {prec}")
    s, prim, bfld = P.orszag_tang(32)
    g = s.ng
    nt = s.ntot
    # active cells (i, j) -> full-grid (i+g, j+g); linear index j*Ntot_x + i
    prim[1][0, g + 7, g + 10] = np.nan     # lowest index of the three
    prim[4][0, g + 12, g + 5] = -0.1
    prim[0][0, g + 20, g + 20] = -1.0
    ctx = Context(32, 32, 1, precision=prec)
    ctx.upload(prim, bfld)
    ctx.init_state()
    d = ctx.diag()
    ctx.close()
    want = (g + 7) * nt[0] + (g + 10)
    say(f"  bad cell {d['bad_cell']} {d['bad_ijk']} kind {d['bad_kind']}, nonfinite "
        f"{d['nonfinite']}, p<=0 {d['npneg']}, rho<=0 {d['nrhoneg']}, step {d['step']}")
    assert d["bad_cell"] == want
    assert d["bad_ijk"] == (g + 10, g + 7, 0)
    assert d["bad_kind"] == 1
    assert d["nonfinite"] == 1 and d["npneg"] == 1 and d["nrhoneg"] == 1
    assert d["pmin"] == pytest.approx(-0.1) and d["rhomin"] == pytest.approx(-1.0)
    assert d["step"] == 0 and d["time"] == 0.0


def test_diag_reports_no_failure_and_the_step_on_a_healthy_run():
    from bonzo_amd._lib import Context
    say("case: diag after 3 healthy steps of turbulence 16^3")
    s, prim, bfld = P.turbulence(16)
    ctx = Context(16, 16, 16)
    ctx.upload(prim, bfld)
    ctx.init_state()
    ctx.run_steps(3)
    d = ctx.diag()
    t, n, _ = ctx.time()
    ctx.close()
    assert d["bad_cell"] == -1 and d["bad_ijk"] == (-1, -1, -1) and d["bad_kind"] == 0
    assert d["nonfinite"] == 0 and d["npneg"] == 0 and d["nrhoneg"] == 0
    assert d["step"] == n == 3 and d["time"] == t > 0


def test_diag_kind_bits_of_a_pressure_failure():
    from bonzo_amd._lib import Context
    say("case: diag of a 1D tube with p<=0 and rho<=0 in one cell")
    rng = np.random.Generator(np.random.PCG64(3))
    s, prim, bfld = P.riemann_tube(64, rng)
    g = s.ng
    prim[4][0, 0, g + 40] = 0.0
    prim[0][0, 0, g + 40] = -2.0
    ctx = Context(64, bc=(1,) * 6)
    ctx.upload(prim, bfld)
    ctx.init_state()
    d = ctx.diag()
    ctx.close()
    assert d["bad_cell"] == g + 40 and d["bad_kind"] == 2 | 4
    assert d["nonfinite"] == 0 and d["npneg"] == 1 and d["nrhoneg"] == 1
