"""bnz_rsolver_rows_noalloc, the allocation-free variant of the
Riemann-solver seam (bnz_hll_flux / bnz_hlld_flux / bnz_hllt_flux): on the
same host rows it writes what the seam writes, bit for bit, on two calls in a
row (the second with fewer faces, so the fixed staging still holds the first
call's longer rows), and a pencil longer than its fixed staging is refused
with BNZ_ERANGE before any device work, leaving the rows as they were."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENTINEL = 7.25


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _rows(a):
    """double** over the rows of a C-contiguous 2D float64 array."""
    return (_dp * a.shape[0])(*[ctypes.cast(a[q].ctypes.data, _dp) for q in range(a.shape[0])])


def _rows_in(rng, n):
    from tests import mathref as M
    wl, wr, bx = M.random_states(rng, n, bx_min=0.0)
    wl[1] += rng.uniform(-3, 3, n)
    wr[1] += rng.uniform(-3, 3, n)
    wl[:, 7] = np.nan  # NaN on both sides: no region matches, the row stays
    wr[:, 7] = np.nan
    return wl, wr, bx


def _seam(L, rs, wl, wr, bx, i1, i2):
    got = np.full((9, wl.shape[1]), SENTINEL)
    getattr(L, f"bnz_{rs}_flux")(_rows(got), _rows(wl), _rows(wr), _ptr(bx), i1, i2, 5.0 / 3.0)
    return got


def _noalloc(L, rs, wl, wr, bx, i1, i2):
    from bonzo_amd._lib import RS
    got = np.full((9, wl.shape[1]), SENTINEL)
    rc = L.bnz_rsolver_rows_noalloc(RS[rs], _rows(got), _rows(wl), _rows(wr), _ptr(bx), i1, i2,
                                    5.0 / 3.0)
    return rc, got


def test_noalloc_seam_equals_the_seam_on_two_calls():
    from bonzo_amd._lib import BNZ_NOALLOC_MAX_FACES, BNZ_OK, load
    L = load()
    rng = np.random.default_rng(77)
    # first call: the longest pencil the fixed staging takes; second: fewer faces
    for call, (n, i1, i2) in enumerate(((BNZ_NOALLOC_MAX_FACES + 8, 3,
                                         BNZ_NOALLOC_MAX_FACES + 2),
                                        (700, 2, 601))):
        wl, wr, bx = _rows_in(rng, n)
        for rs in ("hll", "hlld", "hllt"):
            say(f"case: noalloc seam call {call + 1}, {rs}, {i2 - i1 + 1} faces (i1={i1})")
            want = _seam(L, rs, wl, wr, bx, i1, i2)
            rc, got = _noalloc(L, rs, wl, wr, bx, i1, i2)
            assert rc == BNZ_OK
            assert np.array_equal(got, want, equal_nan=True), rs
            assert np.all(got[:, :i1] == SENTINEL) and np.all(got[:, i2 + 1:] == SENTINEL)
            if rs != "hll":
                assert np.all(got[:, 7] == SENTINEL)


def test_noalloc_seam_refuses_a_pencil_beyond_its_staging():
    from bonzo_amd._lib import BNZ_ERANGE, BNZ_NOALLOC_MAX_FACES, RS, load
    L = load()
    n = BNZ_NOALLOC_MAX_FACES + 1
    wl, wr, bx = _rows_in(np.random.default_rng(5), n)
    got = np.full((9, n), SENTINEL)
    say(f"case: noalloc seam, {n} faces (one beyond the staging)")
    rc = L.bnz_rsolver_rows_noalloc(RS["hlld"], _rows(got), _rows(wl), _rows(wr), _ptr(bx), 0,
                                    n - 1, 5.0 / 3.0)
    assert rc == BNZ_ERANGE
    assert np.all(got == SENTINEL)
