"""ctypes binding of the C ABI in include/tetemoko_hip.h (libtetemoko_hip.so).

The shared library is built in-tree (tetemoko_amd/_lib) by build.py; loading
fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libtetemoko_hip.so")

NCOMP = 22
NBDRY_RS = 11
NBDRY_SW = 9
MAXPATCH = 8
GHOST = 4

OK = 0
E_PSI_NAN = -1
E_OVERFLOW = -10
E_BAD_BC = -2
E_PRED_ORDER = -3
E_NO_BDRY = -4
E_BOX = -5
E_ARG = -6
E_HIP = -7
E_UNSUPPORTED = -8
E_NOMEM = -9
E_SCRATCH = -14

PREC_FP64 = 0
PREC_FP32_BULK = 1

EXPORTED = (
    "tmk_ctx_create", "tmk_ctx_destroy", "tmk_strerror", "tmk_device_ok", "tmk_mem_info",
    "tmk_cache_info", "tmk_cache_set_limit", "tmk_cache_trim", "tmk_batch_done", "tmk_amr_done",
    "tmk_level_set_tile_order",
    "tmk_level_footprint", "tmk_level_footprint_params", "tmk_batch_create", "tmk_batch_destroy",
    "tmk_batch_add", "tmk_batch_remove", "tmk_batch_run", "tmk_batch_sync", "tmk_level_fill",
    "tmk_level_download_fault", "tmk_level_tag", "tmk_patch_update_state", "tmk_patch_plastic_update",
    "tmk_level_create", "tmk_level_destroy", "tmk_level_upload", "tmk_level_download", "tmk_level_step",
    "tmk_level_run", "tmk_level_run_timed", "tmk_level_profile", "tmk_level_sync", "tmk_set_tile",
    "tmk_level_profile_kernels", "tmk_batch_member_error", "tmk_stations_create", "tmk_stations_read",
    "tmk_stations_destroy", "tmk_amr_create", "tmk_amr_destroy", "tmk_amr_run", "tmk_amr_run_timed",
    "tmk_amr_sync", "tmk_make_boxes", "tmk_level_regrid", "tmk_amr_batch_create", "tmk_amr_add",
    "tmk_amr_remove", "tmk_amr_advance", "tmk_amr_member_error", "tmk_patch_update_state_noalloc",
    "tmk_set_march", "tmk_level_set_update_shape", "tmk_level_get_update_shape", "tmk_level_autotune",
)

TUNE_MAX = 32


class Profile(ctypes.Structure):
    _fields_ = [("ms_step", ctypes.c_double), ("ms_interior", ctypes.c_double), ("ms_boundary", ctypes.c_double),
                ("cells_interior", ctypes.c_longlong), ("cells_boundary", ctypes.c_longlong),
                ("tiles_interior", ctypes.c_int), ("tiles_boundary", ctypes.c_int), ("in_graph", ctypes.c_int)]


class Tuning(ctypes.Structure):
    _fields_ = [("march_rows", ctypes.c_int), ("bx", ctypes.c_int), ("by", ctypes.c_int),
                ("ms_per_step", ctypes.c_double), ("ntried", ctypes.c_int),
                ("march_rows_tried", ctypes.c_int * TUNE_MAX), ("bx_tried", ctypes.c_int * TUNE_MAX),
                ("by_tried", ctypes.c_int * TUNE_MAX), ("ms_tried", ctypes.c_double * TUNE_MAX),
                ("xcd", ctypes.c_int), ("xcd_tried", ctypes.c_int * TUNE_MAX)]


class MeshParams(ctypes.Structure):
    _fields_ = [("fill_ratio", ctypes.c_double), ("block_factor", ctypes.c_int), ("max_grid_size", ctypes.c_int),
                ("ref_ratio", ctypes.c_int), ("nest_buffer", ctypes.c_int)]


class Box(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_int * 2), ("hi", ctypes.c_int * 2)]

    @staticmethod
    def make(lo0, lo1, hi0, hi1):
        b = Box()
        b.lo[0], b.lo[1], b.hi[0], b.hi[1] = lo0, lo1, hi0, hi1
        return b

    def tup(self):
        return (self.lo[0], self.lo[1], self.hi[0], self.hi[1])


_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [
        ("cs", _D), ("cp", _D), ("G", _D),
        ("sxx0", _D), ("syy0", _D), ("szz0", _D), ("sxy0", _D), ("sxz0", _D), ("syz0", _D),
        ("use_plasticity", ctypes.c_int),
        ("p_mu", _D), ("p_beta", _D), ("p_eta", _D), ("p_b", _D), ("p_h", _D),
        ("bc", ctypes.c_int * 4),
        ("fault_kind", ctypes.c_int),
        ("r0", _D), ("nuc_time", _D), ("nsig", _D), ("rs_a", _D), ("rs_b", _D), ("rs_V0", _D), ("rs_f0", _D),
        ("rs_L", _D), ("rs_fw", _D), ("rs_Vw", _D), ("rs_fexp", _D),
        ("x0", _D), ("y0", _D), ("rs_rupture_thr", _D), ("ramp_x", _D), ("ramp_w", _D), ("ramp_a", _D),
        ("ramp_Vw", _D),
        ("rs_psi0", _D),
        ("state_law", ctypes.c_int),
        ("fs", _D), ("fd", _D), ("dw", _D), ("sm_value", _D), ("sw_rupture_thr", _D),
        ("num_patches", ctypes.c_int),
        ("xc", _D * MAXPATCH), ("xw", _D * MAXPATCH), ("zc", _D * MAXPATCH), ("zw", _D * MAXPATCH),
        ("tau_p", _D * MAXPATCH),
        ("fric_center", _D * 2), ("fric_width", _D * 2), ("outside_f", _D),
        ("face_lo", ctypes.c_int), ("nface", ctypes.c_int),
        ("f_a", _PD), ("f_b", _PD), ("f_L", _PD), ("f_fs", _PD), ("f_fd", _PD), ("f_dc", _PD),
        ("f_syy", _PD), ("f_tau", _PD),
        ("normal_pred", ctypes.c_int), ("fourth_order", ctypes.c_int), ("prim_limiting", ctypes.c_int),
        ("char_limiting", ctypes.c_int),
        ("dx", _D),
        ("domain", Box),
    ]


class TagParams(ctypes.Structure):
    _fields_ = [("refine_thresh", _D), ("plastic_thresh", _D), ("tag_buffer", ctypes.c_int),
                ("x_coarsen", _D * 2), ("y_coarsen", _D * 2), ("slope_coarsen", _D), ("width_coarsen", _D),
                ("x_coarsen2", _D * 2), ("y_coarsen2", _D * 2), ("level", ctypes.c_int),
                ("coarsen2_max_level", ctypes.c_int), ("stress_thresh", _D), ("sliprate_thresh", _D)]

    @staticmethod
    def make(refine_thresh=0.01, plastic_thresh=1e-5, tag_buffer=2, x_coarsen=(-1e40, 1e40),
             y_coarsen=(-1e40, 1e40), slope_coarsen=0.0, width_coarsen=1e40, x_coarsen2=(-1e40, 1e40),
             y_coarsen2=(-1e40, 1e40), level=0, coarsen2_max_level=3, stress_thresh=0.0, sliprate_thresh=0.0):
        """Defaults of amrGodunov.cpp:263-450."""
        t = TagParams()
        t.refine_thresh, t.plastic_thresh, t.tag_buffer = refine_thresh, plastic_thresh, tag_buffer
        t.x_coarsen[:] = x_coarsen
        t.y_coarsen[:] = y_coarsen
        t.slope_coarsen, t.width_coarsen = slope_coarsen, width_coarsen
        t.x_coarsen2[:] = x_coarsen2
        t.y_coarsen2[:] = y_coarsen2
        t.level, t.coarsen2_max_level = level, coarsen2_max_level
        t.stress_thresh, t.sliprate_thresh = stress_thresh, sliprate_thresh
        return t


class TmkError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__("%s: tetemoko_hip error %d (%s)" % (where, code, strerror(code)))


_lib = None


def lib():
    """Load libtetemoko_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtetemoko_hip.so is not built (%s); run python -c "
                               "'import __graft_entry__; __graft_entry__.build()'" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        c_int, PBox = ctypes.c_int, ctypes.POINTER(Box)
        vp = ctypes.c_void_p
        L.tmk_ctx_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params), c_int]
        L.tmk_ctx_destroy.argtypes = [vp]
        L.tmk_strerror.argtypes = [c_int]
        L.tmk_strerror.restype = ctypes.c_char_p
        L.tmk_device_ok.argtypes = []
        L.tmk_patch_update_state.argtypes = [vp, Box, _PD, Box, c_int, _PD, _PD, _PD, Box, c_int, _D, _D, _PD]
        L.tmk_patch_update_state_noalloc.argtypes = L.tmk_patch_update_state.argtypes
        L.tmk_patch_plastic_update.argtypes = [vp, Box, _PD, Box, c_int, _D]
        L.tmk_level_create.argtypes = [vp, c_int, PBox, ctypes.POINTER(vp)]
        L.tmk_level_destroy.argtypes = [vp]
        L.tmk_level_upload.argtypes = [vp, c_int, _PD, _PD]
        L.tmk_level_download.argtypes = [vp, c_int, _PD, _PD]
        L.tmk_level_step.argtypes = [vp, _D, _D, _PD]
        L.tmk_level_run.argtypes = [vp, c_int, _D, _D]
        L.tmk_level_run_timed.argtypes = [vp, c_int, _D, _D, _PD]
        L.tmk_level_profile.argtypes = [vp, c_int, _D, _D, _PD]
        L.tmk_level_profile_kernels.argtypes = [vp, c_int, _D, _D, ctypes.POINTER(Profile)]
        L.tmk_level_sync.argtypes = [vp]
        L.tmk_set_tile.argtypes = [vp, c_int, c_int]
        L.tmk_set_march.argtypes = [vp, c_int]
        L.tmk_level_set_update_shape.argtypes = [vp, c_int, c_int, c_int]
        L.tmk_level_get_update_shape.argtypes = [vp, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                                 ctypes.POINTER(c_int)]
        L.tmk_level_autotune.argtypes = [vp, _D, _D, c_int, ctypes.POINTER(Tuning)]
        L.tmk_level_set_tile_order.argtypes = [vp, c_int]
        sz = ctypes.c_size_t
        L.tmk_mem_info.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.tmk_cache_info.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.tmk_cache_set_limit.argtypes = [sz]
        L.tmk_cache_trim.argtypes = []
        L.tmk_batch_done.argtypes = [vp, ctypes.POINTER(c_int)]
        L.tmk_amr_done.argtypes = [vp, ctypes.POINTER(c_int)]
        L.tmk_level_footprint.argtypes = [vp, c_int, PBox, ctypes.POINTER(sz)]
        L.tmk_level_footprint_params.argtypes = [ctypes.POINTER(Params), c_int, c_int, PBox, ctypes.POINTER(sz)]
        L.tmk_batch_create.argtypes = [ctypes.POINTER(vp)]
        L.tmk_batch_destroy.argtypes = [vp]
        L.tmk_batch_add.argtypes = [vp, vp, _D, _D]
        L.tmk_batch_remove.argtypes = [vp, vp]
        L.tmk_batch_run.argtypes = [vp, c_int]
        L.tmk_batch_sync.argtypes = [vp]
        L.tmk_batch_member_error.argtypes = [vp, vp, ctypes.POINTER(c_int)]
        L.tmk_stations_create.argtypes = [vp, c_int, _PD, c_int, ctypes.POINTER(vp)]
        L.tmk_stations_read.argtypes = [vp, _PD, c_int, ctypes.POINTER(c_int)]
        L.tmk_stations_destroy.argtypes = [vp]
        L.tmk_level_fill.argtypes = [vp, c_int, _PD, _PD]
        L.tmk_level_download_fault.argtypes = [vp, c_int, _PD]
        L.tmk_level_tag.argtypes = [vp, ctypes.POINTER(TagParams), ctypes.POINTER(ctypes.c_ubyte)]
        L.tmk_amr_create.argtypes = [c_int, ctypes.POINTER(vp), c_int, ctypes.POINTER(vp)]
        L.tmk_amr_destroy.argtypes = [vp]
        L.tmk_amr_run.argtypes = [vp, c_int, _D, _D]
        L.tmk_amr_run_timed.argtypes = [vp, c_int, _D, _D, _PD]
        L.tmk_amr_sync.argtypes = [vp]
        L.tmk_make_boxes.argtypes = [c_int, c_int, ctypes.POINTER(ctypes.c_ubyte), c_int, PBox,
                                     ctypes.POINTER(MeshParams), c_int, PBox, ctypes.POINTER(c_int)]
        L.tmk_level_regrid.argtypes = [vp, vp, vp, c_int]
        L.tmk_amr_batch_create.argtypes = [c_int, c_int, ctypes.POINTER(vp)]
        L.tmk_amr_add.argtypes = [vp, ctypes.POINTER(vp), _D, _D]
        L.tmk_amr_remove.argtypes = [vp, vp]
        L.tmk_amr_advance.argtypes = [vp, c_int]
        L.tmk_amr_member_error.argtypes = [vp, vp, ctypes.POINTER(c_int)]
        _lib = L
    return _lib


def strerror(code):
    try:
        return lib().tmk_strerror(code).decode()
    except RuntimeError:
        return "library not loaded"


def _check(code, where):
    if code != 0:
        raise TmkError(code, where)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_PD)


def make_params(cfg):
    """tmk_params from a scenario config dict (see tetemoko_amd.scenarios); keeps face arrays alive."""
    p = Params()
    keep = []
    p.cs, p.cp, p.G = cfg["cs"], cfg["cp"], cfg["G"]
    bg = cfg.get("background", (0.0,) * 6)
    p.sxx0, p.syy0, p.szz0, p.sxy0, p.sxz0, p.syz0 = bg
    p.use_plasticity = int(cfg.get("plasticity", 0))
    p.p_mu = cfg.get("p_mu", 0.5735)
    p.p_beta = cfg.get("p_beta", p.p_mu / 2)
    p.p_eta = cfg.get("p_eta", 0.2775)
    p.p_b = cfg.get("p_b", 0.0)
    p.p_h = cfg.get("p_h", 0.0)
    for k, v in enumerate(cfg.get("bc", (0, 0, 0, 0))):
        p.bc[k] = v
    p.fault_kind = 0 if cfg.get("fault", "rs") == "rs" else 1
    rs = cfg.get("rs", {})
    for name, key, dflt in (("r0", "nuc_R", 0.0), ("nuc_time", "nuc_T", 0.0), ("nsig", "nuc_S", 0.0),
                            ("rs_a", "a", 0.01), ("rs_b", "b", 0.014), ("rs_V0", "V0", 1e-6),
                            ("rs_f0", "f0", 0.6), ("rs_L", "L", 0.02), ("rs_fw", "fw", 0.13),
                            ("rs_Vw", "Vw", 0.17), ("rs_fexp", "fric_exp", -1.0), ("x0", "nuc_x", 0.0),
                            ("y0", "nuc_y", 0.0), ("rs_rupture_thr", "rupture_thr", 0.001),
                            ("ramp_x", "ramp_x", 0.0), ("ramp_w", "ramp_w", 1e40),
                            ("ramp_a", "ramp_a", 0.0), ("ramp_Vw", "ramp_Vw", 0.0), ("rs_psi0", "psi", 0.6)):
        setattr(p, name, float(rs.get(key, dflt)))
    p.state_law = int(rs.get("state_law", 0))
    sw = cfg.get("sw", {})
    p.fs = sw.get("f_static", 0.677)
    p.fd = sw.get("f_dynamic", 0.525)
    p.dw = sw.get("d_weak", 0.4)
    p.sm_value = sw.get("smooth", 0.0)
    p.sw_rupture_thr = sw.get("rupture_thr", 0.001)
    patches = sw.get("patches", [])
    p.num_patches = len(patches)
    for k, (xc, xw, tau) in enumerate(patches):
        p.xc[k], p.xw[k], p.tau_p[k] = xc, xw, tau
    p.fric_center[0], p.fric_center[1] = sw.get("fric_center", (0.0, 0.0))
    p.fric_width[0], p.fric_width[1] = sw.get("fric_width", (1e40, 1e40))
    p.outside_f = sw.get("outside_f", 1e4)
    faces = cfg.get("faces", {})
    p.face_lo = int(faces.get("lo", 0))
    p.nface = 0
    for name in ("a", "b", "L", "fs", "fd", "dc", "syy", "tau"):
        arr = faces.get(name)
        if arr is not None:
            arr = np.ascontiguousarray(arr, dtype=np.float64)
            keep.append(arr)
            p.nface = len(arr)
            setattr(p, "f_" + name, arr.ctypes.data_as(_PD))
    num = cfg.get("numerics", {})
    p.normal_pred = {"CTU": 0, "PLM": 1, "PPM": 2}[num.get("normal_predictor", "PPM")]
    p.fourth_order = int(num.get("fourth_order", 1))
    p.prim_limiting = int(num.get("prim_limiting", 1))
    p.char_limiting = int(num.get("char_limiting", 0))
    p.dx = cfg["dx"]
    p.domain = Box.make(0, 0, cfg["nx"] - 1, cfg["ny"] - 1)
    return p, keep


class Context:
    """One scenario: parameters (the reference's COMMON blocks), stream, device scratch."""

    def __init__(self, cfg, precision=PREC_FP64):
        self.cfg = cfg
        self.params, self._keep = make_params(cfg)
        h = ctypes.c_void_p()
        _check(lib().tmk_ctx_create(ctypes.byref(h), ctypes.byref(self.params), precision), "tmk_ctx_create")
        self.h = h
        self.precision = precision

    def close(self):
        if self.h:
            lib().tmk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, bx, by):
        _check(lib().tmk_set_tile(self.h, bx, by), "tmk_set_tile")

    def set_march(self, rows):
        """Row-marching interior update in strips of 58 x rows (0: the tile kernels)."""
        _check(lib().tmk_set_march(self.h, rows), "tmk_set_march")

    def update_state(self, valid, U, ubox, B=None, bbox=None, F0=None, F1=None, time=0.0, dt=1.0, noalloc=False):
        """LinElastPatchGodunov::updateState on host FArrayBox arrays (in place). Returns max wave speed.
        noalloc: through tmk_patch_update_state_noalloc (no device allocation; fixed per-device scratch)."""
        mws = ctypes.c_double()
        bb = Box.make(*bbox) if bbox is not None else Box.make(0, 0, -1, -1)
        nb = B.shape[0] if B is not None else 0
        fn = lib().tmk_patch_update_state_noalloc if noalloc else lib().tmk_patch_update_state
        code = fn(self.h, Box.make(*valid), _ptr(U), Box.make(*ubox), NCOMP,
                  _ptr(F0), _ptr(F1), _ptr(B), bb, nb, time, dt, ctypes.byref(mws))
        _check(code, "tmk_patch_update_state_noalloc" if noalloc else "tmk_patch_update_state")
        return mws.value

    def plastic_update(self, valid, U, ubox, dt):
        _check(lib().tmk_patch_plastic_update(self.h, Box.make(*valid), _ptr(U), Box.make(*ubox), NCOMP, dt),
               "tmk_patch_plastic_update")


class Level:
    """Device-resident level of boxes (LinElastLevelGodunov)."""

    def __init__(self, ctx, boxes):
        self.ctx = ctx
        self.boxes = [tuple(b) for b in boxes]
        arr = (Box * len(boxes))(*[Box.make(*b) for b in boxes])
        h = ctypes.c_void_p()
        _check(lib().tmk_level_create(ctx.h, len(boxes), arr, ctypes.byref(h)), "tmk_level_create")
        self.h = h

    def close(self):
        if self.h:
            lib().tmk_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, ibox, U, B=None):
        _check(lib().tmk_level_upload(self.h, ibox, _ptr(U), _ptr(B)), "tmk_level_upload")

    def download(self, ibox, U, B=None):
        _check(lib().tmk_level_download(self.h, ibox, _ptr(U), _ptr(B)), "tmk_level_download")

    def fill(self, ibox, state22, B=None):
        """Uniform state on the device (no host state array) plus the box's fault data."""
        st = np.ascontiguousarray(state22, dtype=np.float64)
        _check(lib().tmk_level_fill(self.h, ibox, _ptr(st), _ptr(B)), "tmk_level_fill")

    def tag(self, tp, nx, ny):
        """Refinement tags of the current state, (ny, nx) uint8."""
        out = np.zeros((ny, nx), dtype=np.uint8)
        _check(lib().tmk_level_tag(self.h, ctypes.byref(tp), out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))),
               "tmk_level_tag")
        return out

    def download_fault(self, ibox, B):
        _check(lib().tmk_level_download_fault(self.h, ibox, _ptr(B)), "tmk_level_download_fault")

    def step(self, time, dt):
        dtn = ctypes.c_double()
        _check(lib().tmk_level_step(self.h, time, dt, ctypes.byref(dtn)), "tmk_level_step")
        return dtn.value

    def run(self, nsteps, time, dt):
        _check(lib().tmk_level_run(self.h, nsteps, time, dt), "tmk_level_run")

    def run_timed(self, nsteps, time, dt):
        """Run nsteps; returns elapsed ms from HIP events on the level's stream."""
        ms = ctypes.c_double()
        _check(lib().tmk_level_run_timed(self.h, nsteps, time, dt, ctypes.byref(ms)), "tmk_level_run_timed")
        return ms.value

    def profile(self, nsteps, time, dt):
        """Mean ms per eager step (one fused launch per step for a single-box level)."""
        ms = ctypes.c_double()
        _check(lib().tmk_level_profile(self.h, nsteps, time, dt, ctypes.byref(ms)), "tmk_level_profile")
        return ms.value

    def profile_kernels(self, nsteps, time, dt):
        """Per-launch means (ms) of the interior-tile and boundary-tile update kernels and of
        the whole step, each timed with HIP events on the stream it runs on, plus the cells
        and tiles each launch covers."""
        p = Profile()
        _check(lib().tmk_level_profile_kernels(self.h, nsteps, time, dt, ctypes.byref(p)),
               "tmk_level_profile_kernels")
        return {k: getattr(p, k) for k, _ in Profile._fields_}

    def sync(self):
        _check(lib().tmk_level_sync(self.h), "tmk_level_sync")

    def set_update_shape(self, march_rows, bx=0, by=0):
        """The level's own update shape: row-marching strips of march_rows rows (0: tiles of
        bx x by; 0, 0 = the default tile)."""
        _check(lib().tmk_level_set_update_shape(self.h, march_rows, bx, by), "tmk_level_set_update_shape")

    def update_shape(self):
        """(march_rows, bx, by) the update runs (march_rows 0: the tile kernels)."""
        r, x, y = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().tmk_level_get_update_shape(self.h, ctypes.byref(r), ctypes.byref(x), ctypes.byref(y)),
               "tmk_level_get_update_shape")
        return r.value, x.value, y.value

    def set_tile_order(self, xcd):
        """Interior workgroups in XCD bands (1, the default) or row-major order (0)."""
        _check(lib().tmk_level_set_tile_order(self.h, int(xcd)), "tmk_level_set_tile_order")

    def autotune(self, time, dt, nsteps=10):
        """Times every update shape from the level's current state, then the fastest with the
        other workgroup order; sets the fastest and returns the level to that state:
        {"march_rows", "bx", "by", "xcd", "ms_per_step", "tried": [(rows, bx, by, xcd, ms), ...]}."""
        t = Tuning()
        _check(lib().tmk_level_autotune(self.h, time, dt, nsteps, ctypes.byref(t)), "tmk_level_autotune")
        tried = [(t.march_rows_tried[k], t.bx_tried[k], t.by_tried[k], t.xcd_tried[k], t.ms_tried[k])
                 for k in range(t.ntried)]
        return {"march_rows": t.march_rows, "bx": t.bx, "by": t.by, "xcd": t.xcd, "ms_per_step": t.ms_per_step,
                "tried": tried}


class Amr:
    """Hierarchy of device-resident levels stepped together (subcycled, ratio r)."""

    def __init__(self, levels, r):
        self.levels = list(levels)
        arr = (ctypes.c_void_p * len(levels))(*[lv.h for lv in levels])
        h = ctypes.c_void_p()
        _check(lib().tmk_amr_create(len(levels), arr, r, ctypes.byref(h)), "tmk_amr_create")
        self.h = h
        self.r = r

    def close(self):
        if self.h:
            lib().tmk_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps, time, dt0):
        _check(lib().tmk_amr_run(self.h, nsteps, time, dt0), "tmk_amr_run")

    def run_timed(self, nsteps, time, dt0):
        ms = ctypes.c_double()
        _check(lib().tmk_amr_run_timed(self.h, nsteps, time, dt0, ctypes.byref(ms)), "tmk_amr_run_timed")
        return ms.value

    def sync(self):
        _check(lib().tmk_amr_sync(self.h), "tmk_amr_sync")


class AmrBatch:
    """Hierarchies of several scenarios (same level count and ratio) in shared launches."""

    def __init__(self, nlevels, r):
        h = ctypes.c_void_p()
        _check(lib().tmk_amr_batch_create(nlevels, r, ctypes.byref(h)), "tmk_amr_batch_create")
        self.h = h
        self.nlevels, self.r = nlevels, r

    def close(self):
        if self.h:
            lib().tmk_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, levels, time, dt0):
        arr = (ctypes.c_void_p * len(levels))(*[lv.h for lv in levels])
        _check(lib().tmk_amr_add(self.h, arr, time, dt0), "tmk_amr_add")

    def remove(self, level0):
        _check(lib().tmk_amr_remove(self.h, level0.h), "tmk_amr_remove")

    def advance(self, nsteps):
        _check(lib().tmk_amr_advance(self.h, nsteps), "tmk_amr_advance")

    def sync(self):
        _check(lib().tmk_amr_sync(self.h), "tmk_amr_sync")

    def done(self):
        """True when every step enqueued so far has completed (does not wait)."""
        d = ctypes.c_int()
        _check(lib().tmk_amr_done(self.h, ctypes.byref(d)), "tmk_amr_done")
        return bool(d.value)

    def member_error(self, level0):
        code = ctypes.c_int()
        _check(lib().tmk_amr_member_error(self.h, level0.h, ctypes.byref(code)), "tmk_amr_member_error")
        return code.value


def make_boxes(tags, coarse_boxes=None, ref_ratio=4, block_factor=8, max_grid_size=64, fill_ratio=0.75,
               nest_buffer=0):
    """Boxes of the next finer level (finer-level cells) from a level's tags (ny, nx) uint8."""
    tags = np.ascontiguousarray(tags, dtype=np.uint8)
    ny, nx = tags.shape
    mp = MeshParams(fill_ratio, block_factor, max_grid_size, ref_ratio, nest_buffer)
    cb = coarse_boxes or []
    carr = (Box * max(len(cb), 1))(*[Box.make(*b) for b in cb])
    cap = 1024
    while True:
        out = (Box * cap)()
        n = ctypes.c_int()
        code = lib().tmk_make_boxes(nx, ny, tags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), len(cb), carr,
                                    ctypes.byref(mp), cap, out, ctypes.byref(n))
        if code == E_OVERFLOW:
            cap = n.value
            continue
        _check(code, "tmk_make_boxes")
        return [out[k].tup() for k in range(n.value)]


def level_regrid(fresh, old, coarse, ref_ratio):
    """The new layout's data: the old layout's cells where it covers them, else interpolated
    from the coarser level (tmk_level_regrid)."""
    _check(lib().tmk_level_regrid(fresh.h, old.h if old is not None else None, coarse.h, ref_ratio),
           "tmk_level_regrid")


def mem_info():
    """(free, total) device bytes."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().tmk_mem_info(ctypes.byref(f), ctypes.byref(t)), "tmk_mem_info")
    return f.value, t.value


def cache_info():
    """(idle cached, in use, limit) bytes of the level-array cache on the current device."""
    c, u, m = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().tmk_cache_info(ctypes.byref(c), ctypes.byref(u), ctypes.byref(m)), "tmk_cache_info")
    return c.value, u.value, m.value


def cache_set_limit(nbytes):
    """Limit on the level-array cache's idle bytes per device (0: no caching)."""
    _check(lib().tmk_cache_set_limit(int(nbytes)), "tmk_cache_set_limit")


def cache_trim():
    """Free every idle block of the level-array cache on the current device."""
    _check(lib().tmk_cache_trim(), "tmk_cache_trim")


def level_footprint(ctx, boxes):
    arr = (Box * len(boxes))(*[Box.make(*b) for b in boxes])
    n = ctypes.c_size_t()
    _check(lib().tmk_level_footprint(ctx.h, len(boxes), arr, ctypes.byref(n)), "tmk_level_footprint")
    return n.value


def level_footprint_params(cfg, boxes, precision=PREC_FP64):
    """Device bytes of a level of these boxes for a scenario config, without a context."""
    p, keep = make_params(cfg)
    arr = (Box * len(boxes))(*[Box.make(*b) for b in boxes])
    n = ctypes.c_size_t()
    _check(lib().tmk_level_footprint_params(ctypes.byref(p), precision, len(boxes), arr, ctypes.byref(n)),
           "tmk_level_footprint_params")
    del keep
    return n.value


class Batch:
    """Continuous batch: member levels (scenarios) advance together in shared launches."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().tmk_batch_create(ctypes.byref(h)), "tmk_batch_create")
        self.h = h

    def close(self):
        if self.h:
            lib().tmk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, level, time, dt):
        _check(lib().tmk_batch_add(self.h, level.h, time, dt), "tmk_batch_add")

    def remove(self, level):
        _check(lib().tmk_batch_remove(self.h, level.h), "tmk_batch_remove")

    def run(self, nsteps):
        _check(lib().tmk_batch_run(self.h, nsteps), "tmk_batch_run")

    def sync(self):
        _check(lib().tmk_batch_sync(self.h), "tmk_batch_sync")

    def done(self):
        """True when every step enqueued so far has completed (does not wait)."""
        d = ctypes.c_int()
        _check(lib().tmk_batch_done(self.h, ctypes.byref(d)), "tmk_batch_done")
        return bool(d.value)

    def member_error(self, level):
        """Error code the member reported (0 = none); other members are unaffected by it."""
        code = ctypes.c_int()
        _check(lib().tmk_batch_member_error(self.h, level.h, ctypes.byref(code)), "tmk_batch_member_error")
        return code.value


class Stations:
    """Fault stations of a level, recorded on the device after every level step
    (AMRLevelLinElast fault stations, AMRLevelLinElast.cpp:258-309,1758-1810)."""

    def __init__(self, level, x, capacity=1024):
        self.level = level
        self.x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        self.capacity = int(capacity)
        self.nb = NBDRY_RS if level.ctx.cfg["fault"] == "rs" else NBDRY_SW
        h = ctypes.c_void_p()
        _check(lib().tmk_stations_create(level.h, len(self.x), _ptr(self.x), self.capacity, ctypes.byref(h)),
               "tmk_stations_create")
        self.h = h

    def read(self, allow_overflow=False):
        """Records since the last read: array [nrecords, nstations, 1 + nb] (time, fault data)."""
        buf = np.empty((self.capacity, len(self.x), 1 + self.nb))
        n = ctypes.c_int()
        code = lib().tmk_stations_read(self.h, _ptr(buf), self.capacity, ctypes.byref(n))
        if code != 0 and not (allow_overflow and code == E_OVERFLOW):
            _check(code, "tmk_stations_read")
        return buf[:n.value].copy()

    def close(self):
        if self.h:
            lib().tmk_stations_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
