"""CPU-side checks: the C-ABI library loads and exports what include/bnzgpu.h
declares, the host selection API keeps the reference's semantics, and the
initial-condition generators produce solenoidal face fields.  No GPU calls."""
import ctypes
import os
import re
import sys
import subprocess

import numpy as np
import pytest

from bonzo_amd import _lib, problems as P
from bonzo_amd.simulation import BnzIntegr, Reconstr, RSolver, TIntegr, _Params, read_param

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "bnzgpu.h")).read()
    return sorted(set(re.findall(r"\b(bnz_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    syms = header_symbols()
    assert len(syms) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bnz_[a-z0-9_]+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    L = _lib.load()
    for s in syms:
        assert hasattr(L, s)
    assert set(_lib.ABI_SYMBOLS) == set(syms)


def test_params_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of bnz_params / bnz_diag_out agree with the C compiler's
    layout of include/bnzgpu.h."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "bnzgpu.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",'
        ' sizeof(bnz_params),'
        ' offsetof(bnz_params, lmin), offsetof(bnz_params, gam), offsetof(bnz_params, device),'
        ' offsetof(bnz_params, char_proj),'
        ' sizeof(bnz_diag_out), offsetof(bnz_diag_out, divb_max),'
        ' offsetof(bnz_diag_out, nonfinite), offsetof(bnz_diag_out, bad_ijk),'
        ' offsetof(bnz_diag_out, step), offsetof(bnz_diag_out, time),'
        ' sizeof(bnz_box), offsetof(bnz_box, lmin),'
        ' offsetof(bnz_box, tmax)); return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    mine = [ctypes.sizeof(_lib.BnzParams), _lib.BnzParams.lmin.offset, _lib.BnzParams.gam.offset,
            _lib.BnzParams.device.offset, _lib.BnzParams.char_proj.offset,
            ctypes.sizeof(_lib.BnzDiag),
            _lib.BnzDiag.divb_max.offset, _lib.BnzDiag.nonfinite.offset,
            _lib.BnzDiag.bad_ijk.offset, _lib.BnzDiag.step.offset, _lib.BnzDiag.time.offset,
            ctypes.sizeof(_lib.BnzBox), _lib.BnzBox.lmin.offset, _lib.BnzBox.tmax.offset]
    assert vals == mine


def test_create_rejects_bad_selection_without_touching_gpu():
    L = _lib.load()
    p = _lib.BnzParams()
    p.nx, p.ny, p.nz = 8, 8, 8
    p.rsolver = _lib.RS["hllc"]
    p.reconstr = 1
    p.gam, p.cour = 5.0 / 3.0, 0.4
    h = ctypes.c_void_p()
    rc = L.bnz_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == _lib.BNZ_EUNSUPPORTED
    assert b"hydro-only" in L.bnz_last_error(h)
    L.bnz_destroy(h)
    p.rsolver = _lib.RS["hlld"]
    p.gam = 1.0
    rc = L.bnz_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == _lib.BNZ_EINVAL
    L.bnz_destroy(h)
    # fewer active cells than ghost layers on a periodic or reflecting axis:
    # the reference's copies would read unfilled ghosts; outflow is fine
    p.gam = 5.0 / 3.0
    p.nx, p.ny, p.nz = 2, 1, 1
    for flags, ok in (((0,) * 6, False), ((2, 2, 0, 0, 0, 0), False), ((1,) * 6, True)):
        for q in range(6):
            p.bc[q] = flags[q]
        rc = L.bnz_create(ctypes.byref(p), ctypes.byref(h))
        if ok:
            assert rc in (_lib.BNZ_OK, _lib.BNZ_EHIP)   # validation passes (no GPU here)
        else:
            assert rc == _lib.BNZ_EINVAL
            assert b"at least ng" in L.bnz_last_error(h)
        L.bnz_destroy(h)


def test_batch_create_rejects_bad_selection_without_touching_gpu():
    L = _lib.load()
    p = _lib.make_params(precision="fp32c", rsolver="hllt")
    boxes = (_lib.BnzBox * 2)()
    for b in boxes:
        b.nx, b.ny, b.nz, b.tmax = 16, 1, 1, 0.1
        for a in range(3):
            b.lmax[a] = 1.0
    h = ctypes.c_void_p()
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 2, ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    assert b"FP64" in L.bnz_batch_last_error(h)
    L.bnz_batch_destroy(h)
    p = _lib.make_params(rsolver="hllc")
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 2, ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    L.bnz_batch_destroy(h)
    p = _lib.make_params()
    boxes[1].bc[0] = 3  # user boundaries are single-grid only
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 2, ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    assert b"box 1" in L.bnz_batch_last_error(h)
    L.bnz_batch_destroy(h)
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 0, ctypes.byref(h)) == _lib.BNZ_EINVAL
    # oversized boxes are rejected before any host or device array is sized
    boxes[1].bc[0] = 0
    for size in ((2 ** 30 + 1, 1, 1), (2 ** 31 - 1, 1, 1), (1300, 1300, 1300)):
        boxes[1].nx, boxes[1].ny, boxes[1].nz = size
        assert L.bnz_batch_create(ctypes.byref(p), boxes, 2, ctypes.byref(h)) == _lib.BNZ_EINVAL
        assert b"too large" in L.bnz_batch_last_error(h)
        L.bnz_batch_destroy(h)


def test_create_rejects_oversized_grids_without_touching_gpu():
    L = _lib.load()
    h = ctypes.c_void_p()
    for size in ((2 ** 30 + 1, 1, 1), (2 ** 31 - 1, 1, 1), (1300, 1300, 1300)):
        p = _lib.make_params(nx=size[0], ny=size[1], nz=size[2])
        assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EINVAL
        assert b"too large" in L.bnz_last_error(h)
        L.bnz_destroy(h)


def test_selection_semantics_follow_reference(tmp_path):
    cfg = tmp_path / "config.cfg"
    cfg.write_text("[computation]\nNx: 16\nNy: 1\nNz: 1\ntmax: 0.1\ncour: 0.4\n"
                   "tintegr: nonsense\nrsolver: weird\nreconstr: foo\ncharact_proj: 0\n"
                   "[physics]\ngam: 1.4\n")
    assert read_param("computation", "Nx", "i", str(tmp_path)) == 16
    integr = BnzIntegr(_Params(str(tmp_path)))
    # unknown names fall back to VL / HLL / linear (integrator.pyx:220-242)
    assert integr.rsolver == RSolver.RS_HLL
    assert integr.reconstr == Reconstr.RCN_LINEAR and integr.reconstr_order == 1
    with pytest.raises(ValueError):
        BnzIntegr(_Params(config={"computation": {"cour": 0.4, "rsolver": "hllc"},
                                  "physics": {"gam": 1.4}}))
    # rk3 selects the RK3 integrator, whose grid has ng = 9 (grid.pyx:142-143)
    it = BnzIntegr(_Params(config={"computation": {"cour": 0.4, "tintegr": "rk3"},
                                   "physics": {"gam": 1.4}}))
    assert it.tintegr == TIntegr.TINT_RK3
    from bonzo_amd.simulation import BnzGrid
    cfg = {"computation": {"cour": 0.4, "tintegr": "rk3", "Nx": 16, "Ny": 16, "Nz": 1},
           "physics": {"gam": 1.4}}
    assert BnzGrid(_Params(config=cfg), BnzIntegr(_Params(config=cfg))).ng == 9


@pytest.mark.parametrize("key", ["gx", "gy", "gz", "f0", "kappa0", "nuiic0", "kL", "mu", "eta",
                                 "mu4", "eta4"])
def test_unbuilt_physics_is_refused_not_dropped(key):
    """Gravity, turbulence driving and diffusion act in the reference as soon
    as their coefficient is non-zero (integrator.pyx:146-190,
    diffusion.pyx:158-164); they are not built here, so such a configuration
    raises instead of silently running without them.  Zero values (the
    reference's sample config) are accepted."""
    base = {"computation": {"cour": 0.4, "rsolver": "hlld"}, "physics": {"gam": 1.4}}
    zero = {k: dict(v) for k, v in base.items()}
    zero["physics"][key] = 0.0
    BnzIntegr(_Params(config=zero))
    bad = {k: dict(v) for k, v in base.items()}
    bad["physics"][key] = 0.5
    with pytest.raises(NotImplementedError, match=key):
        BnzIntegr(_Params(config=bad))


def test_config_dict_with_usr_dir_that_has_no_config_file(tmp_path):
    """A dict of sections with a usr_dir that holds no config.cfg (the usr_dir
    only receives out/ and rst/): keys the dict lacks take their defaults.
    Without a dict the missing file still raises."""
    from bonzo_amd.simulation import BnzGrid
    cfg = {"computation": {"Nx": 16, "Ny": 16, "Nz": 1, "tmax": 0.1, "cour": 0.4,
                           "rsolver": "hlld"},
           "physics": {"gam": 5.0 / 3.0}}
    p = _Params(str(tmp_path), cfg)
    integr = BnzIntegr(p)
    assert integr.tintegr == TIntegr.TINT_VL and integr.rsolver == RSolver.RS_HLLD
    assert tuple(BnzGrid(p, integr).Nact) == (16, 16, 1)
    with pytest.raises(FileNotFoundError):
        BnzIntegr(_Params(str(tmp_path)))
    # with the file present, the dict still overrides it key by key
    (tmp_path / "config.cfg").write_text("[computation]\ncour: 0.2\ntintegr: rk3\n")
    p = _Params(str(tmp_path), cfg)
    integr = BnzIntegr(p)
    assert integr.cour == 0.4 and integr.tintegr == TIntegr.TINT_RK3


def test_non_cartesian_geometry_is_refused():
    from bonzo_amd.simulation import BnzGrid
    cfg = {"computation": {"cour": 0.4, "Nx": 16, "Ny": 1, "Nz": 1},
           "physics": {"gam": 1.4, "geometry": "cyl"}}
    integr = BnzIntegr(_Params(config=cfg))
    with pytest.raises(NotImplementedError, match="geometry"):
        BnzGrid(_Params(config=cfg), integr)
    cfg["physics"]["geometry"] = "car"
    BnzGrid(_Params(config=cfg), integr)


def _divb(s, bfld):
    g = s.ng
    nx, ny, nz = s.nact
    sk, sj, si = s.active()
    d = (bfld[0][sk, sj, g + 1:g + nx + 1] - bfld[0][sk, sj, si]) / s.dx(0)
    if s.dims >= 2:
        d = d + (np.roll(bfld[1][sk, sj, si], -1, axis=1) - bfld[1][sk, sj, si]) / s.dx(1)
    if s.dims >= 3:
        d = d + (np.roll(bfld[2][sk, sj, si], -1, axis=0) - bfld[2][sk, sj, si]) / s.dx(2)
    return d


def test_turbulence_ic_is_solenoidal_and_scaled():
    s, prim, bfld = P.turbulence(32)
    sk, sj, si = s.active()
    # periodic wrap of the x faces for the check
    b = bfld.copy()
    b[0][sk, sj, s.ng + 32] = b[0][sk, sj, s.ng]
    assert np.abs(_divb(s, b)).max() < 1e-11
    v = prim[1:4][:, sk, sj, si]
    assert abs(np.sqrt(np.mean(np.sum(v ** 2, axis=0))) - 0.5) < 1e-12
    # solenoidal velocity: spectral divergence vanishes
    vk = [np.fft.fftn(v[c]) for c in range(3)]
    k = np.fft.fftfreq(32, d=1.0 / 32)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    div = KX * vk[0] + KY * vk[1] + KZ * vk[2]
    assert np.abs(div).max() < 1e-9 * np.abs(vk[0]).max()
    assert np.allclose(prim[0][sk, sj, si], 1.0) and np.allclose(prim[4][sk, sj, si], 0.6)


def test_orszag_tang_ic_is_solenoidal():
    s, prim, bfld = P.orszag_tang(64)
    sk, sj, si = s.active()
    b = bfld.copy()
    b[0][sk, sj, s.ng + 64] = b[0][sk, sj, s.ng]
    assert np.abs(_divb(s, b)).max() < 1e-12


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU's share; --gpus N > 1 exits with a message
    before touching any device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_kernel_accounting():
    """Units and algorithmic bytes per launch of the bench's roofline figure:
    flux3 over the union of the face ranges, updates over every cell."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    n = 256
    nt = n + 6
    m = (nt - 2 - 1) - 2 + 1          # PLM face range [2, nt - 3] per axis
    assert bench.kernel_units("flux3_plm", n) == (m + 2) ** 3
    assert bench.kernel_units("update_corr", n) == nt ** 3
    assert bench.kernel_bytes("fp64", "update_corr", n) == 45 * 8 * nt ** 3
    faces = 3 * (m + 1) * (m + 2) ** 2
    assert bench.kernel_bytes("fp64", "flux3_plm", n) == (m + 2) ** 3 * 12 * 8 + faces * 8 * 8


def test_product_fails_loudly_without_the_library(monkeypatch):
    """No CPU fallback: without libbnzgpu.so the package raises instead of
    computing anything, and nothing under bonzo_amd/ imports the oracle."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libbnzgpu.so")
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for dirpath, _, files in os.walk(os.path.join(root, "bonzo_amd")):
        for f in files:
            if f.endswith(".py"):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f


def test_noalloc_seam_rejects_bad_calls_without_touching_gpu():
    """bnz_rsolver_rows_noalloc checks its arguments and its fixed size
    before any device call: a pencil longer than BNZ_NOALLOC_MAX_FACES gives
    BNZ_ERANGE, an unknown solver or i2 < i1 BNZ_EINVAL, and the rows stay
    as they were."""
    L = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    n = _lib.BNZ_NOALLOC_MAX_FACES + 1
    bufs = [np.full((9, n), 7.25) for _ in range(3)]
    bx = np.zeros(n)
    rows = [(dp * 9)(*[ctypes.cast(b[q].ctypes.data, dp) for q in range(9)]) for b in bufs]
    bxp = bx.ctypes.data_as(dp)
    gam = 5.0 / 3.0
    assert L.bnz_rsolver_rows_noalloc(_lib.RS["hlld"], *rows, bxp, 0, n - 1, gam) == \
        _lib.BNZ_ERANGE
    assert L.bnz_rsolver_rows_noalloc(_lib.RS["hllc"], *rows, bxp, 0, 10, gam) == \
        _lib.BNZ_EINVAL
    assert L.bnz_rsolver_rows_noalloc(_lib.RS["hll"], *rows, bxp, 10, 9, gam) == \
        _lib.BNZ_EINVAL
    assert L.bnz_rsolver_rows_noalloc(_lib.RS["hll"], *rows, bxp, 0, 10, 1.0) == \
        _lib.BNZ_EINVAL
    assert all(np.all(b == 7.25) for b in bufs)


def test_hydro_build_selection_follows_reference():
    """mfield=False is the reference's hydro build: hll and hllc are bound,
    hlld and hlla are not (integrator.pyx:253-268); hllt and characteristic
    limiting of the hydro build are not built here and raise."""
    base = {"computation": {"cour": 0.4, "rsolver": "hllc"}, "physics": {"gam": 1.4}}
    it = BnzIntegr(_Params(config=base), mfield=False)
    assert it.rsolver == RSolver.RS_HLLC and not it.mfield
    with pytest.raises(ValueError):
        BnzIntegr(_Params(config=base))            # hllc with the field: NULL in the reference
    for r, exc in (("hlld", ValueError), ("hlla", ValueError), ("hllt", NotImplementedError)):
        cfg = {"computation": {"cour": 0.4, "rsolver": r}, "physics": {"gam": 1.4}}
        with pytest.raises(exc):
            BnzIntegr(_Params(config=cfg), mfield=False)
    cfg = {"computation": {"cour": 0.4, "rsolver": "hll", "charact_proj": 1},
           "physics": {"gam": 1.4}}
    with pytest.raises(NotImplementedError):
        BnzIntegr(_Params(config=cfg), mfield=False)


def test_create_validates_the_hydro_build_without_touching_gpu():
    L = _lib.load()
    h = ctypes.c_void_p()
    bad = [dict(rsolver="hlld", hydro=True), dict(rsolver="hllt", hydro=True),
           dict(rsolver="hllc", hydro=True, precision="fp32c"),
           dict(rsolver="hllc", hydro=True, char_proj=True)]
    for kw in bad:
        p = _lib.make_params(nx=16, ny=16, **kw)
        assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED, kw
        L.bnz_destroy(h)
    p = _lib.make_params(nx=16, ny=16, rsolver="hllc")
    assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    assert b"hydro" in L.bnz_last_error(h)
    L.bnz_destroy(h)
    p = _lib.make_params(nx=16, ny=16, rsolver="hll")
    p.hydro = 2
    assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EINVAL
    L.bnz_destroy(h)
    boxes = (_lib.BnzBox * 1)()
    boxes[0].nx, boxes[0].ny, boxes[0].nz, boxes[0].tmax = 16, 1, 1, 0.1
    for a in range(3):
        boxes[0].lmax[a] = 1.0
    p = _lib.make_params(rsolver="hll", hydro=True)
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 1, ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    L.bnz_batch_destroy(h)


def test_upwind_emf_selection_and_validation_without_touching_gpu():
    """[computation] ct_upwind picks the upwind corner EMF (interp_ee2); the
    device takes it in FP64 on single grids and refuses FP32C and batches."""
    cfg = {"computation": {"cour": 0.4, "rsolver": "hlld", "ct_upwind": 1},
           "physics": {"gam": 1.4}}
    assert BnzIntegr(_Params(config=cfg)).ct_upwind == 1
    cfg["computation"]["ct_upwind"] = 2
    with pytest.raises(ValueError):
        BnzIntegr(_Params(config=cfg))
    L = _lib.load()
    h = ctypes.c_void_p()
    p = _lib.make_params(nx=16, ny=16, ct_upwind=True, precision="fp32c")
    assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    L.bnz_destroy(h)
    p = _lib.make_params(nx=16, ny=16)
    p.ct_upwind = 3
    assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EINVAL
    L.bnz_destroy(h)
    boxes = (_lib.BnzBox * 1)()
    boxes[0].nx, boxes[0].ny, boxes[0].nz, boxes[0].tmax = 16, 16, 1, 0.1
    for a in range(3):
        boxes[0].lmax[a] = 1.0
    p = _lib.make_params(ct_upwind=True)
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 1, ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    L.bnz_batch_destroy(h)


def test_rk3_validation_without_touching_gpu():
    L = _lib.load()
    h = ctypes.c_void_p()
    p = _lib.make_params(nx=16, ny=16, tintegr="rk3", precision="fp32c")
    assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    L.bnz_destroy(h)
    p = _lib.make_params(nx=8, ny=8, tintegr="rk3")  # periodic axes need ng = 9 cells
    assert L.bnz_create(ctypes.byref(p), ctypes.byref(h)) == _lib.BNZ_EINVAL
    assert b"at least ng" in L.bnz_last_error(h)
    L.bnz_destroy(h)
    boxes = (_lib.BnzBox * 1)()
    boxes[0].nx, boxes[0].ny, boxes[0].nz, boxes[0].tmax = 16, 1, 1, 0.1
    for a in range(3):
        boxes[0].lmax[a] = 1.0
    p = _lib.make_params(tintegr="rk3")
    assert L.bnz_batch_create(ctypes.byref(p), boxes, 1, ctypes.byref(h)) == _lib.BNZ_EUNSUPPORTED
    L.bnz_batch_destroy(h)


def test_user_bc_specs_on_the_host():
    """user_bc specs become the slab arrays of bnz_set_user_bc: slab order
    x fastest over ng layers times the full other extents; Fixed over the
    active rows leaves the others alone (-2); Mapped gives flat source
    indices and refuses sources off the grid; flag 3 without a user
    boundary for the problem fails before any device work."""
    from bonzo_amd.simulation import Simulation
    from bonzo_amd.user_bc import Fixed, Mapped, slab, side_arrays
    ntot, ng = (10, 8, 1), 3
    i, j, k = slab(ntot, ng, 0, 1)
    assert i.size == ng * 8 and list(i[:3]) == [7, 8, 9] and list(j[:4]) == [0, 0, 0, 1]
    i, j, k = slab(ntot, ng, 1, 0)
    assert i.size == 10 * ng and list(i[:10]) == list(range(10)) and j.max() == ng - 1
    w = np.arange(9.0)
    src, fixed = side_arrays(Fixed(w, rows="active"), ntot, ng, 2, 0, 0)
    jj = slab(ntot, ng, 0, 0)[1]
    act = (jj >= ng) & (jj < 8 - ng)
    assert np.all(src[act] == -1) and np.all(src[~act] == -2)
    assert fixed.shape == (src.size, 12) and np.array_equal(fixed[0, :9], w)
    assert np.array_equal(fixed[0, 9:], w[6:9])
    src, fixed = side_arrays(Mapped(lambda i, j, k: (i, j + 2, k)), ntot, ng, 2, 1, 0)
    i, j, k = slab(ntot, ng, 1, 0)
    assert fixed is None and np.array_equal(src, (j + 2) * 10 + i)
    with pytest.raises(ValueError, match="off the grid|out of the grid"):
        side_arrays(Mapped(lambda i, j, k: (i - 5, j, k)), ntot, ng, 2, 1, 0)
    with pytest.raises(ValueError, match="9 values"):
        Fixed(np.zeros(8))
    cfg = {"computation": {"Nx": 16, "Ny": 8, "Nz": 1, "tmax": 0.1, "cour": 0.4,
                           "rsolver": "hlld", "reconstr": "linear"},
           "physics": {"gam": 5.0 / 3.0, "bc_x1": 3}, "problem": {"name": "otvortex"}}
    with pytest.raises(ValueError, match="defines no user boundaries"):
        Simulation(config=cfg)
    cfg["physics"]["bc_x1"] = 4
    with pytest.raises(ValueError, match="boundary flag 4"):
        Simulation(config=cfg)


def test_oblique_tube_setup():
    """The st2d setup: Ryu-Jones 2a states rotated onto (1, 2)/sqrt 5, the
    state and face fields exactly invariant under (i, j) -> (i - 2, j + 1),
    discrete div B zero to round-off, and the two far-field face fields the
    rotated tube fields."""
    from bonzo_amd import problems as P
    s, prim, bfld, user = P.oblique_tube(8, aspect=4)
    assert (s.nx, s.ny) == (32, 8) and s.dx(0) == s.dx(1)
    for arr in (prim, bfld):
        for a in arr:
            assert np.array_equal(a[0, :-1, 2:], a[0, 1:, :-2])
    g = s.ng
    b = bfld[:, 0]
    div = (b[0][g:-g, g + 1:-g + 1] - b[0][g:-g, g:-g]) + (b[1][g + 1:-g + 1, g:-g] - b[1][g:-g, g:-g])
    assert np.abs(div).max() < 1e-14
    c4 = 1 / np.sqrt(4 * np.pi)
    r5 = np.sqrt(5.0)
    bl = np.array([(2 - 2 * 3.6) / r5, (4 + 3.6) / r5]) * c4
    br = np.array([(2 - 8) / r5, (4 + 4) / r5]) * c4
    assert np.allclose([b[0][g, g], b[1][g, g]], bl, rtol=1e-13)
    assert np.allclose([b[0][-g - 1, -g - 1], b[1][-g - 1, -g - 1]], br, rtol=1e-13)
    assert np.isclose(prim[1, 0, g, g], (1.2 - 2 * 0.01) / r5) and prim[0, 0, -g - 1, -g - 1] == 1.0
    assert set(user) == {(0, 0), (0, 1), (1, 0), (1, 1)}
