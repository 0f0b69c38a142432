"""The single-grid work-shape autotuner (bonzo_amd/autotune.py) against a
stub context whose kernel times are known: it must pick each class's
fastest shape even when the device reads slow during its first timed runs,
which is when the first candidate used to be measured."""
from bonzo_amd.autotune import CANDIDATES, autotune


class StubContext:
    """Per-class milliseconds per step for each shape (the shape of the
    flux class decides the flux time, and so on); timed runs among the
    first ``slow_runs`` read ``slow`` times longer."""

    def __init__(self, times, ntot=(40, 40, 40), slow_runs=3, slow=1.08):
        self.times = times
        self.ntot = ntot
        self.shapes = (0, 0, 0)
        self.slow_runs = slow_runs
        self.slow = slow
        self.timed_runs = 0
        self.timing = False
        self.acc = {}

    def set_tuning(self, flux, emf, update):
        self.shapes = (flux, emf, update)

    def run_steps(self, n):
        if not self.timing:
            return
        f = self.slow if self.timed_runs < self.slow_runs else 1.0
        self.timed_runs += 1
        for name, cls, shape in (("flux3_plm", "flux", self.shapes[0]),
                                 ("emf_corr", "emf", self.shapes[1]),
                                 ("update_corr", "update", self.shapes[2])):
            tot, cnt = self.acc.get(name, (0.0, 0))
            self.acc[name] = (tot + n * f * self.times[shape][cls], cnt + n)

    def synchronize(self):
        pass

    def set_kernel_timing(self, on):
        if on:
            self.acc = {}
        self.timing = on

    def kernel_timing(self):
        return dict(self.acc)


def _times(best):
    # every shape 3 % slower than the best of its class
    return {c: {k: (1.0 if c == best[k] else 1.03) for k in ("flux", "emf", "update")}
            for c in CANDIDATES}


def test_picks_the_fastest_shape_per_class():
    ctx = StubContext(_times({"flux": 16, "emf": 0, "update": 8}), slow_runs=0)
    best, table = autotune(ctx)
    assert best == (16, 0, 8)
    assert ctx.shapes == (16, 0, 8)
    assert set(table) == set(CANDIDATES)


def test_first_candidate_is_not_penalised_by_a_slow_start():
    """The first candidate (flat, 0) is the fastest, but the device reads
    8 % slow for the first three timed runs: a single pass in candidate
    order would pick another shape; two passes, the second reversed, keep
    the flat one."""
    ctx = StubContext(_times({"flux": 0, "emf": 0, "update": 0}), slow_runs=3)
    best, _ = autotune(ctx)
    assert best == (0, 0, 0)
    one_pass = StubContext(_times({"flux": 0, "emf": 0, "update": 0}), slow_runs=3)
    assert autotune(one_pass, passes=1)[0] != (0, 0, 0)  # the bias this guards against


def test_grids_without_work_shapes_keep_the_defaults():
    ctx = StubContext(_times({"flux": 16, "emf": 16, "update": 16}), ntot=(40, 40, 1))
    best, table = autotune(ctx)
    assert best == (0, 0, 0) and table == {}
    assert ctx.timed_runs == 0
