"""FP32-compensated path against the FP64 path (which the FP64 parity tests
tie to the oracle), through the C ABI.

Bars (stated for the 3D turbulence workload): relative L1 <= 1e-5 per
variable against FP64 after N steps on the same initial state, and
max |div B| <= 1e-12 in code units.  Shock tubes and the Orszag-Tang vortex
are checked at looser, stated tolerances because FP32 rounding differences
move discontinuities by a fraction of a cell.
"""
import sys

import numpy as np
import pytest

from bonzo_amd import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l1(a, b):
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


def pair(prob, steps, bc=(0,) * 6, tmax=None):
    from bonzo_amd._lib import Context
    s, prim, bfld = prob
    out = {}
    for prec in ("fp64", "fp32c"):
        g = Context(s.nx, s.ny, s.nz, bc=bc, precision=prec)
        g.upload(prim, bfld)
        g.init_state()
        d0 = g.diag()
        if tmax is None:
            g.run_steps(steps)
        else:
            g.advance(tmax)
        p_, b_ = g.download()
        out[prec] = dict(prim=p_, bfld=b_, diag=g.diag(), diag0=d0, time=g.time())
        g.close()
    return s, out


def errors(s, out):
    sk, sj, si = s.active()
    e = [rel_l1(out["fp32c"]["prim"][v][sk, sj, si], out["fp64"]["prim"][v][sk, sj, si])
         for v in range(9)]
    e += [rel_l1(out["fp32c"]["bfld"][c][sk, sj, si], out["fp64"]["bfld"][c][sk, sj, si])
          for c in range(3)]
    return e


def test_turbulence_fp32c_accuracy_and_divb():
    say("case: turbulence 64^3, 40 steps, fp32c vs fp64")
    s, out = pair(P.turbulence(64, seed_b=71, seed_v=72), 40)
    e = errors(s, out)
    say("  rel L1 per variable: " + " ".join(f"{x:.2e}" for x in e))
    d = out["fp32c"]["diag"]
    say(f"  divB fp32c {d['divb_max']:.3e}  fp64 {out['fp64']['diag']['divb_max']:.3e}  "
        f"guards {d['hlld_guards']}")
    assert max(e[v] for v in (0, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11)) <= 1e-5
    assert d["divb_max"] <= 1e-12
    d0 = out["fp32c"]["diag0"]
    assert abs(d["mass"] - d0["mass"]) <= 1e-12 * abs(d0["mass"])
    assert abs(d["energy"] - d0["energy"]) <= 1e-12 * abs(d0["energy"])


def test_orszag_tang_fp32c():
    say("case: Orszag-Tang 128^2 to t=0.2, fp32c vs fp64")
    s, out = pair(P.orszag_tang(128), None, tmax=0.2)
    e = errors(s, out)
    say("  rel L1 per variable: " + " ".join(f"{x:.2e}" for x in e))
    assert out["fp32c"]["time"][0] == pytest.approx(0.2, rel=1e-12)
    assert max(e[v] for v in (0, 4, 6, 7, 9, 10)) <= 1e-4
    assert out["fp32c"]["diag"]["divb_max"] <= 1e-11


def test_brio_wu_fp32c():
    say("case: Brio-Wu 1D 512 cells, outflow, to t=0.1, fp32c vs fp64")
    s, out = pair(P.brio_wu(512, axis=0, transverse=0), None, bc=(1,) * 6, tmax=0.1)
    e = errors(s, out)
    say("  rel L1 per variable: " + " ".join(f"{x:.2e}" for x in e))
    assert max(e[v] for v in (0, 1, 4, 7)) <= 1e-4


def test_fp32c_init_and_download_round_trip():
    """Upload, init and download without steps returns the FP64 state."""
    from bonzo_amd._lib import Context
    say("case: fp32c upload/init/download round trip, turbulence 24^3")
    s, prim, bfld = P.turbulence(24, seed_b=81, seed_v=82)
    g64 = Context(24, 24, 24)
    g64.upload(prim, bfld)
    g64.init_state()
    p64, b64 = g64.download()
    g = Context(24, 24, 24, precision="fp32c")
    g.upload(prim, bfld)
    g.init_state()
    p32, b32 = g.download()
    assert np.array_equal(b32, b64)
    assert np.max(np.abs(p32 - p64)) <= 1e-13 * max(1.0, np.max(np.abs(p64)))


def test_reflecting_walls_fp32c():
    """FP32C ghost fill with reflecting walls (sign flips on the hi words and
    the FP32 prim, zeroed wall faces; the lo words are not filled, since only
    active cells read them) against the FP64 path on a smooth 2D state."""
    from tests.test_oracle_pins import _wall_state
    say("case: smooth 2D state 64x48, reflecting x walls, 40 steps, fp32c vs fp64")
    s, out = pair(_wall_state(64, 48, 0), 40, bc=(2, 2, 0, 0, 0, 0))
    e = errors(s, out)
    say("  rel L1 per variable: " + " ".join(f"{x:.2e}" for x in e))
    assert max(e[v] for v in (0, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11)) <= 1e-5
    assert out["fp32c"]["diag"]["divb_max"] <= 1e-12
    g = s.ng
    assert np.all(out["fp32c"]["bfld"][0][0, g:g + 48, g] == 0)


def test_fp32c_work_shapes_do_not_change_results():
    """Flat, tile (runs flat) and slab-column shapes of the FP32C flux, EMF
    and update kernels give the same state bit for bit: each cell's
    arithmetic does not depend on the shape, and the CFL maximum does not
    depend on the reduction order."""
    from bonzo_amd._lib import Context
    say("case: FP32C turbulence 20^3 and 2D OT 32^2, shapes (0,0,0) (8,0,16) (1,1,1) (5,7,3), 4 steps")
    for s, prim, bfld in (P.turbulence(20, seed_b=63, seed_v=64), P.orszag_tang(32)):
        res = []
        for shapes in ((0, 0, 0), (8, 0, 16), (1, 1, 1), (5, 7, 3)):
            g = Context(s.nx, s.ny, s.nz, precision="fp32c")
            g.set_tuning(*shapes)
            g.upload(prim, bfld)
            g.init_state()
            g.run_steps(4)
            res.append(g.save_state())
            g.close()
        for r in res[1:]:
            assert np.array_equal(r, res[0])


def test_brio_wu_walls_1d_fp32c():
    """1D FP32C steps form their EMFs in the update (zero on a reflecting
    wall plane) instead of an EMF kernel: a tube between reflecting walls
    against the FP64 path."""
    say("case: 1D tube 400 cells, reflecting walls, no normal field, 150 steps, fp32c vs fp64")
    rng = np.random.Generator(np.random.PCG64(808))
    s, prim, bfld = P.riemann_tube(400, rng)
    prim[6][:] = 0.0  # no field through the walls
    bfld[0][:] = 0.0
    s, out = pair((s, prim, bfld), 150, bc=(2, 2, 0, 0, 0, 0))
    e = errors(s, out)
    say("  rel L1 per variable: " + " ".join(f"{x:.2e}" for x in e))
    assert max(e[v] for v in (0, 1, 4, 7, 8, 10, 11)) <= 1e-5
    assert out["fp32c"]["diag"]["nonfinite"] == 0
