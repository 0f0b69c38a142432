"""ctypes binding of libbnzgpu.so (declarations in include/bnzgpu.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C bonzo_amd/csrc``).  There is no fallback: if the shared object is
missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbnzgpu.so")

BNZ_OK, BNZ_EINVAL, BNZ_ENOMEM, BNZ_EHIP, BNZ_EUNSUPPORTED, BNZ_ESTATE = 0, -1, -2, -3, -4, -5
BNZ_ERANGE = -6
BNZ_NOALLOC_MAX_FACES = 16384
RS = {"hll": 0, "hllc": 1, "hlld": 2, "hlla": 3, "hllt": 4}
RCN = {"const": 0, "linear": 1, "weno": 2, "parab": 3}
PREC = {"fp64": 0, "fp32c": 1, "fp64fma": 2}

ABI_SYMBOLS = (
    "bnz_create", "bnz_destroy", "bnz_last_error", "bnz_get_shape", "bnz_upload",
    "bnz_download", "bnz_init_state", "bnz_fill_ghosts", "bnz_new_dt", "bnz_vl_step",
    "bnz_advance", "bnz_run_steps", "bnz_get_time", "bnz_set_time", "bnz_diag", "bnz_history",
    "bnz_synchronize", "bnz_set_tuning", "bnz_set_user_bc", "bnz_set_step_graph",
    "bnz_set_kernel_timing",
    "bnz_kernel_timing",
    "bnz_state_size", "bnz_save_state", "bnz_load_state",
    "bnz_gpu_flux_pencils", "bnz_hll_flux", "bnz_hlld_flux", "bnz_hllt_flux", "bnz_hllc_flux",
    "bnz_rsolver_rows_noalloc",
    "bnz_batch_create", "bnz_batch_destroy", "bnz_batch_last_error", "bnz_batch_box_shape",
    "bnz_batch_upload", "bnz_batch_download", "bnz_batch_init_state", "bnz_batch_advance",
    "bnz_batch_get_times", "bnz_batch_synchronize", "bnz_batch_set_tuning",
)


class BnzParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("rsolver", ctypes.c_int), ("reconstr", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("lmin", ctypes.c_double * 3), ("lmax", ctypes.c_double * 3),
        ("gam", ctypes.c_double), ("cour", ctypes.c_double),
        ("dt_fms_ref", ctypes.c_int), ("precision", ctypes.c_int), ("device", ctypes.c_int),
        ("char_proj", ctypes.c_int), ("hydro", ctypes.c_int), ("ct_upwind", ctypes.c_int),
        ("tintegr", ctypes.c_int),
    ]


class BnzDiag(ctypes.Structure):
    _fields_ = [
        ("mass", ctypes.c_double), ("mom", ctypes.c_double * 3), ("energy", ctypes.c_double),
        ("scalar", ctypes.c_double), ("divb_max", ctypes.c_double), ("bmax", ctypes.c_double),
        ("pmin", ctypes.c_double), ("rhomin", ctypes.c_double), ("nonfinite", ctypes.c_long),
        ("hlld_guards", ctypes.c_long), ("npneg", ctypes.c_long), ("nrhoneg", ctypes.c_long),
        ("bad_cell", ctypes.c_long), ("bad_ijk", ctypes.c_int * 3), ("bad_kind", ctypes.c_int),
        ("step", ctypes.c_int), ("time", ctypes.c_double),
    ]


class BnzBox(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("lmin", ctypes.c_double * 3), ("lmax", ctypes.c_double * 3),
        ("tmax", ctypes.c_double),
    ]


class BnzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bnz error {code}: {msg}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load():
    """Load libbnzgpu.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "or make -C bonzo_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.bnz_create.argtypes = [ctypes.POINTER(BnzParams), ctypes.POINTER(vp)]
    L.bnz_destroy.argtypes = [vp]
    L.bnz_last_error.restype = ctypes.c_char_p
    L.bnz_last_error.argtypes = [vp]
    L.bnz_get_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.bnz_upload.argtypes = [vp, _dp, _dp]
    L.bnz_download.argtypes = [vp, _dp, _dp, _dp]
    for name in ("bnz_init_state", "bnz_fill_ghosts", "bnz_synchronize"):
        getattr(L, name).argtypes = [vp]
    L.bnz_new_dt.argtypes = [vp, _dp]
    L.bnz_vl_step.argtypes = [vp, ctypes.c_double]
    L.bnz_advance.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.bnz_run_steps.argtypes = [vp, ctypes.c_int]
    L.bnz_get_time.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int), _dp]
    L.bnz_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_int]
    L.bnz_diag.argtypes = [vp, ctypes.POINTER(BnzDiag)]
    L.bnz_history.argtypes = [vp, _dp]
    L.bnz_state_size.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
    L.bnz_save_state.argtypes = [vp, vp, ctypes.c_long]
    L.bnz_load_state.argtypes = [vp, vp, ctypes.c_long]
    L.bnz_set_kernel_timing.argtypes = [vp, ctypes.c_int]
    L.bnz_set_step_graph.argtypes = [vp, ctypes.c_int]
    L.bnz_set_user_bc.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), _dp]
    L.bnz_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.bnz_kernel_timing.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _dp,
                                    ctypes.POINTER(ctypes.c_long)]
    L.bnz_gpu_flux_pencils.argtypes = [ctypes.c_int, vp, vp, vp, vp, ctypes.c_long,
                                       ctypes.c_double]
    for name in ("bnz_hll_flux", "bnz_hlld_flux", "bnz_hllt_flux", "bnz_hllc_flux"):
        getattr(L, name).argtypes = [ctypes.POINTER(_dp)] * 3 + [_dp, ctypes.c_int, ctypes.c_int,
                                                                  ctypes.c_double]
        getattr(L, name).restype = None
    L.bnz_rsolver_rows_noalloc.argtypes = [ctypes.c_int] + [ctypes.POINTER(_dp)] * 3 + [
        _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.bnz_batch_create.argtypes = [ctypes.POINTER(BnzParams), ctypes.POINTER(BnzBox),
                                   ctypes.c_int, ctypes.POINTER(vp)]
    L.bnz_batch_destroy.argtypes = [vp]
    L.bnz_batch_last_error.restype = ctypes.c_char_p
    L.bnz_batch_last_error.argtypes = [vp]
    L.bnz_batch_box_shape.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.bnz_batch_upload.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.bnz_batch_download.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.bnz_batch_init_state.argtypes = [vp]
    L.bnz_batch_synchronize.argtypes = [vp]
    L.bnz_batch_advance.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.bnz_batch_get_times.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int)]
    L.bnz_batch_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    for name in ABI_SYMBOLS:
        if name not in ("bnz_last_error", "bnz_batch_last_error", "bnz_hll_flux",
                        "bnz_hlld_flux", "bnz_hllt_flux", "bnz_hllc_flux"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def make_params(nx=1, ny=1, nz=1, rsolver="hlld", reconstr="linear", bc=(0,) * 6,
                lmin=(0.0, 0.0, 0.0), lmax=(1.0, 1.0, 1.0), gam=5.0 / 3.0, cour=0.4,
                dt_fms_ref=True, precision="fp64", device=0, char_proj=False, hydro=False,
                ct_upwind=False, tintegr="vl"):
    p = BnzParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.rsolver = RS[rsolver] if isinstance(rsolver, str) else int(rsolver)
    p.reconstr = RCN[reconstr] if isinstance(reconstr, str) else int(reconstr)
    for q in range(6):
        p.bc[q] = int(bc[q])
    for q in range(3):
        p.lmin[q] = lmin[q]
        p.lmax[q] = lmax[q]
    p.gam, p.cour = gam, cour
    p.dt_fms_ref = int(bool(dt_fms_ref))
    p.precision = PREC[precision]
    p.device = device
    p.char_proj = int(bool(char_proj))
    p.hydro = int(bool(hydro))
    p.ct_upwind = int(bool(ct_upwind))
    p.tintegr = {"vl": 0, "rk3": 1}[tintegr] if isinstance(tintegr, str) else int(tintegr)
    return p


class Context:
    """One device grid: owns a bnz_ctx (device arrays, stream)."""

    def __init__(self, nx, ny=1, nz=1, rsolver="hlld", reconstr="linear", bc=(0,) * 6,
                 lmin=(0.0, 0.0, 0.0), lmax=(1.0, 1.0, 1.0), gam=5.0 / 3.0, cour=0.4,
                 dt_fms_ref=True, precision="fp64", device=0, char_proj=False, hydro=False,
                 ct_upwind=False, tintegr="vl"):
        L = load()
        p = make_params(nx, ny, nz, rsolver, reconstr, bc, lmin, lmax, gam, cour, dt_fms_ref,
                        precision, device, char_proj, hydro, ct_upwind, tintegr)
        h = ctypes.c_void_p()
        rc = L.bnz_create(ctypes.byref(p), ctypes.byref(h))
        if rc != BNZ_OK:
            msg = L.bnz_last_error(h).decode() if h.value else "bnz_create failed"
            if h.value:
                L.bnz_destroy(h)
            raise BnzError(rc, msg)
        self._h = h
        nt = (ctypes.c_int * 3)()
        ng = ctypes.c_int()
        self._check(L.bnz_get_shape(h, nt, ctypes.byref(ng)))
        self.ntot = (nt[0], nt[1], nt[2])
        self.ng = ng.value

    def _check(self, rc):
        if rc != BNZ_OK:
            raise BnzError(rc, load().bnz_last_error(self._h).decode())

    @property
    def shape(self):
        return (self.ntot[2], self.ntot[1], self.ntot[0])

    def upload(self, prim, bfld):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        bfld = np.ascontiguousarray(bfld, dtype=np.float64)
        assert prim.shape == (9,) + self.shape and bfld.shape == (3,) + self.shape
        self._check(load().bnz_upload(self._h, _ptr(prim), _ptr(bfld)))

    def download(self, cons=False):
        prim = np.empty((9,) + self.shape)
        bfld = np.empty((3,) + self.shape)
        c = np.empty((9,) + self.shape) if cons else None
        self._check(load().bnz_download(self._h, _ptr(prim), _ptr(bfld), _ptr(c)))
        return (prim, bfld, c) if cons else (prim, bfld)

    def init_state(self):
        self._check(load().bnz_init_state(self._h))

    def fill_ghosts(self):
        self._check(load().bnz_fill_ghosts(self._h))

    def new_dt(self):
        d = ctypes.c_double()
        self._check(load().bnz_new_dt(self._h, ctypes.byref(d)))
        return d.value

    def vl_step(self, dt):
        self._check(load().bnz_vl_step(self._h, dt))

    def advance(self, tmax, max_steps=0):
        n = ctypes.c_int()
        self._check(load().bnz_advance(self._h, tmax, max_steps, ctypes.byref(n)))
        return n.value

    def run_steps(self, n):
        self._check(load().bnz_run_steps(self._h, n))

    def synchronize(self):
        self._check(load().bnz_synchronize(self._h))

    def time(self):
        t = ctypes.c_double()
        s = ctypes.c_int()
        d = ctypes.c_double()
        self._check(load().bnz_get_time(self._h, ctypes.byref(t), ctypes.byref(s), ctypes.byref(d)))
        return t.value, s.value, d.value

    def set_time(self, t, step):
        self._check(load().bnz_set_time(self._h, t, step))

    def diag(self):
        d = BnzDiag()
        self._check(load().bnz_diag(self._h, ctypes.byref(d)))
        return dict(mass=d.mass, mom=tuple(d.mom), energy=d.energy, scalar=d.scalar,
                    divb_max=d.divb_max, bmax=d.bmax, pmin=d.pmin, rhomin=d.rhomin,
                    nonfinite=d.nonfinite, hlld_guards=d.hlld_guards, npneg=d.npneg,
                    nrhoneg=d.nrhoneg, bad_cell=d.bad_cell, bad_ijk=tuple(d.bad_ijk),
                    bad_kind=d.bad_kind, step=d.step, time=d.time)

    def history(self):
        """One history row (step, time, 14 cell means; see bnz_history)."""
        out = np.zeros(16)
        self._check(load().bnz_history(self._h, _ptr(out)))
        return out

    def save_state(self):
        """Exact device state (bnz_save_state) as a uint8 array."""
        n = ctypes.c_long()
        self._check(load().bnz_state_size(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        self._check(load().bnz_save_state(self._h, buf.ctypes.data, n.value))
        return buf

    def load_state(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._check(load().bnz_load_state(self._h, buf.ctypes.data, buf.size))

    def set_tuning(self, flux_chunk=0, emf_chunk=0, update_chunk=0):
        self._check(load().bnz_set_tuning(self._h, flux_chunk, emf_chunk, update_chunk))

    def set_user_bc(self, axis, side, spec):
        """User boundary (flag 3) of one side: a user_bc.Fixed or
        user_bc.Mapped spec, or raw (src, fixed) arrays in the layout of
        bnz_set_user_bc."""
        from .user_bc import side_arrays
        dims = 3 if self.ntot[2] > 1 else 2 if self.ntot[1] > 1 else 1
        src, fixed = side_arrays(spec, self.ntot, self.ng, dims, axis, side)
        fp = _ptr(np.ascontiguousarray(fixed)) if fixed is not None else None
        self._check(load().bnz_set_user_bc(
            self._h, axis, side, src.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), fp))

    def set_step_graph(self, on):
        """True: advance() replays captured chunks of eight steps; False (the
        default): it launches every kernel directly (bnz_set_step_graph)."""
        self._check(load().bnz_set_step_graph(self._h, int(bool(on))))

    def set_kernel_timing(self, on):
        self._check(load().bnz_set_kernel_timing(self._h, int(bool(on))))

    def kernel_timing(self):
        out = {}
        q = 0
        name = ctypes.create_string_buffer(64)
        while True:
            tot = ctypes.c_double()
            n = ctypes.c_long()
            rc = load().bnz_kernel_timing(self._h, q, name, 64, ctypes.byref(tot), ctypes.byref(n))
            if rc != BNZ_OK:
                break
            out[name.value.decode()] = (tot.value, n.value)
            q += 1
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().bnz_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent boxes of uneven size advanced together: owns a bnz_batch.

    ``boxes`` is a sequence of dicts with keys nx, ny, nz (active cells),
    bc (6 flags, periodic 0 / outflow 1), lmin, lmax (extent) and tmax.
    The remaining keyword arguments are the common selection of Context.
    """

    def __init__(self, boxes, rsolver="hlld", reconstr="linear", gam=5.0 / 3.0, cour=0.4,
                 dt_fms_ref=True, precision="fp64", device=0):
        L = load()
        p = make_params(rsolver=rsolver, reconstr=reconstr, gam=gam, cour=cour,
                        dt_fms_ref=dt_fms_ref, precision=precision, device=device)
        arr = (BnzBox * len(boxes))()
        for q, b in enumerate(boxes):
            arr[q].nx, arr[q].ny, arr[q].nz = b["nx"], b.get("ny", 1), b.get("nz", 1)
            bc = b.get("bc", (0,) * 6)
            for s in range(6):
                arr[q].bc[s] = int(bc[s])
            for a in range(3):
                arr[q].lmin[a] = b.get("lmin", (0.0, 0.0, 0.0))[a]
                arr[q].lmax[a] = b.get("lmax", (1.0, 1.0, 1.0))[a]
            arr[q].tmax = b["tmax"]
        h = ctypes.c_void_p()
        rc = L.bnz_batch_create(ctypes.byref(p), arr, len(boxes), ctypes.byref(h))
        if rc != BNZ_OK:
            msg = L.bnz_batch_last_error(h).decode() if h.value else "bnz_batch_create failed"
            if h.value:
                L.bnz_batch_destroy(h)
            raise BnzError(rc, msg)
        self._h = h
        self.nbox = len(boxes)
        self.shapes = []
        nt = (ctypes.c_int * 3)()
        for q in range(self.nbox):
            self._check(L.bnz_batch_box_shape(h, q, nt))
            self.shapes.append((nt[2], nt[1], nt[0]))

    def _check(self, rc):
        if rc != BNZ_OK:
            raise BnzError(rc, load().bnz_batch_last_error(self._h).decode())

    def upload(self, box, prim, bfld):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        bfld = np.ascontiguousarray(bfld, dtype=np.float64)
        sh = self.shapes[box]
        assert prim.shape == (9,) + sh and bfld.shape == (3,) + sh
        self._check(load().bnz_batch_upload(self._h, box, _ptr(prim), _ptr(bfld)))

    def download(self, box):
        sh = self.shapes[box]
        prim = np.empty((9,) + sh)
        bfld = np.empty((3,) + sh)
        self._check(load().bnz_batch_download(self._h, box, _ptr(prim), _ptr(bfld)))
        return prim, bfld

    def init_state(self):
        self._check(load().bnz_batch_init_state(self._h))

    def advance(self, max_steps=0, use_graph=True):
        n = ctypes.c_int()
        self._check(load().bnz_batch_advance(self._h, max_steps, int(bool(use_graph)),
                                             ctypes.byref(n)))
        return n.value

    def times(self):
        t = np.empty(self.nbox)
        s = (ctypes.c_int * self.nbox)()
        self._check(load().bnz_batch_get_times(self._h, _ptr(t), s))
        return t, np.array(list(s), dtype=np.int64)

    def synchronize(self):
        self._check(load().bnz_batch_synchronize(self._h))

    def set_tuning(self, flux_blocks=0, emf_blocks=0, update_blocks=0):
        """Workgroup caps of the flux, EMF and update (and fused 1D) launches;
        0 is the default.  Results do not depend on them."""
        self._check(load().bnz_batch_set_tuning(self._h, int(flux_blocks), int(emf_blocks),
                                                int(update_blocks)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().bnz_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
