"""Autotuner for the work shapes of the 3D kernels.

Each candidate shape (0 = one thread per face/edge/cell in x-fastest order,
1 = 64x4 tiles ordered along z, n > 1 = slab columns of n rows walked plane
by plane, so that the next plane's rows are still in the XCD's L2 when they
are reused) is applied to all three kernel classes, a few VL steps are run
with per-launch HIP-event timing on the context's stream, and each class
keeps the candidate with the smallest summed time:
  flux   -- Godunov fluxes of both stages (all directions in one launch)
  emf    -- edge EMF kernels
  update -- stage updates (the corrector carries the fused CFL reduction)
The runs advance the context's state: callers tune first and then upload
their initial state again, or tune a scratch context of the same shape.
"""
CANDIDATES = (0, 1, 4, 8, 16, 32)


def _class_times(kt):
    out = {"flux": 0.0, "emf": 0.0, "update": 0.0}
    for name, (tot, n) in kt.items():
        if name.startswith("flux"):
            out["flux"] += tot
        elif name.startswith("emf"):
            out["emf"] += tot
        elif name.startswith("update") or name.startswith("dt_partial"):
            out["update"] += tot
    return out


def autotune(ctx, steps=2, candidates=CANDIDATES):
    """Pick (flux_chunk, emf_chunk, update_chunk) for ``ctx`` and apply it.

    Returns (best, table) where table[c] holds the per-class milliseconds of
    candidate c.  The context must hold an initialised state."""
    table = {}
    for c in candidates:
        ctx.set_tuning(c, c, c)
        ctx.run_steps(1)                      # first launch of a shape pays setup
        ctx.synchronize()
        ctx.set_kernel_timing(True)
        ctx.run_steps(steps)
        ctx.synchronize()
        ctx.set_kernel_timing(False)
        table[c] = _class_times(ctx.kernel_timing())
    best = tuple(min(candidates, key=lambda c: table[c][k]) for k in ("flux", "emf", "update"))
    ctx.set_tuning(*best)
    return best, table
