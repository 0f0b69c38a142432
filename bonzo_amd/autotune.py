"""Autotuners: work shapes of a single grid's 3D kernels (autotune) and
workgroup caps of a batch's stage launches (autotune_batch).

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
import time

CANDIDATES = (0, 1, 4, 8, 16, 32)
BATCH_CANDIDATES = (2048, 4096, 8192, 16384, 65536)


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


def autotune(ctx, steps=2, candidates=CANDIDATES, passes=2):
    """Pick (flux_chunk, emf_chunk, update_chunk) for ``ctx`` and apply it.

    Returns (best, table) where table[c] holds the per-class milliseconds of
    candidate c, the smallest over ``passes`` measurements.  The candidates
    are measured in order, then in reverse order, and so on, after a
    warm-up of the device: a candidate measured only first, while the clocks
    and caches settle, would read several per cent slow.  The context must
    hold an initialised state.  Work shapes exist on 3D grids only; other
    grids keep the defaults."""
    table = {}
    if ctx.ntot[2] == 1:
        ctx.set_tuning(0, 0, 0)
        return (0, 0, 0), table
    ctx.set_tuning(candidates[0], candidates[0], candidates[0])
    ctx.run_steps(steps)
    for p in range(passes):
        for c in (candidates if p % 2 == 0 else candidates[::-1]):
            ctx.set_tuning(c, c, c)
            ctx.run_steps(1)                      # first launch of a shape pays setup
            ctx.synchronize()
            ctx.set_kernel_timing(True)
            ctx.run_steps(steps)
            ctx.synchronize()
            ctx.set_kernel_timing(False)
            t = _class_times(ctx.kernel_timing())
            table[c] = t if c not in table else {k: min(v, table[c][k]) for k, v in t.items()}
    best = tuple(min(candidates, key=lambda c: table[c][k]) for k in ("flux", "emf", "update"))
    ctx.set_tuning(*best)
    return best, table


def autotune_batch(batch, reset, steps=16, candidates=BATCH_CANDIDATES):
    """Pick the workgroup caps of ``batch``'s flux, EMF and update launches
    (bnz_batch_set_tuning) and apply them.

    The launches walk their work blocks with at most that many workgroups;
    fewer, longer-lived workgroups pay less dispatch, more of them balance
    the end of a launch better.  The classes are tuned one after the other,
    the others held at their best so far; each candidate runs ``steps`` eager
    steps from the state ``reset()`` restores (upload and init_state), timed
    around the synchronising advance.  Returns (best, table) with table[class]
    = {candidate: seconds}; the batch is left reset."""
    best = [0, 0, 0]
    table = {}
    for k, name in enumerate(("flux", "emf", "update")):
        times = {}
        for c in candidates:
            trial = list(best)
            trial[k] = c
            batch.set_tuning(*trial)
            reset()
            batch.advance(max_steps=1, use_graph=False)  # first launch of a size pays setup
            batch.synchronize()
            t0 = time.perf_counter()
            batch.advance(max_steps=steps, use_graph=False)
            times[c] = time.perf_counter() - t0
        best[k] = min(times, key=times.get)
        table[name] = times
    batch.set_tuning(*best)
    reset()
    return tuple(best), table
