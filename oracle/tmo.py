"""ctypes binding of the CPU oracle (oracle/tmo.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(tetemoko_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TMO_LIB names another build of tmo.c (the mutation tests load deliberately broken copies)
LIB_PATH = os.environ.get("TMO_LIB") or os.path.join(_HERE, "_build", "libtmo.so")

NCOMP = 22
NB_RS = 11
NB_SW = 9
MAXPATCH = 8
GHOST = 4


class Box(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_int * 2), ("hi", ctypes.c_int * 2)]

    @staticmethod
    def make(lo0, lo1, hi0, hi1):
        b = Box()
        b.lo[0], b.lo[1], b.hi[0], b.hi[1] = lo0, lo1, hi0, hi1
        return b


class Fab(ctypes.Structure):
    _fields_ = [("d", ctypes.POINTER(ctypes.c_double)), ("b", Box), ("nc", ctypes.c_int)]


_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [
        ("cs", _D), ("cp", _D), ("G", _D), ("rho", _D), ("lam", _D), ("K", _D), ("Zp", _D), ("Zs", _D),
        ("background", _D * NCOMP),
        ("p_mu", _D), ("p_beta", _D), ("p_eta", _D), ("p_b", _D), ("p_h", _D),
        ("use_plasticity", ctypes.c_int),
        ("bc", ctypes.c_int * 4),
        ("fault_kind", ctypes.c_int),
        ("r0", _D), ("nucTime", _D), ("nsig", _D), ("rs_a", _D), ("rs_b", _D), ("rs_V0", _D),
        ("rs_f0", _D), ("rs_L", _D), ("rs_fw", _D), ("rs_Vw", _D), ("rs_fExp", _D),
        ("x0", _D), ("y0", _D), ("rs_rupture_thr", _D), ("ramp_x", _D), ("ramp_w", _D),
        ("ramp_a", _D), ("ramp_Vw", _D),
        ("rs_psi0", _D),
        ("state_law", ctypes.c_int),
        ("fs", _D), ("fd", _D), ("dw", _D), ("smValue", _D), ("sw_rupture_thr", _D),
        ("num_patches", ctypes.c_int),
        ("xc", _D * MAXPATCH), ("xw", _D * MAXPATCH), ("zc", _D * MAXPATCH), ("zw", _D * MAXPATCH),
        ("tau_p", _D * MAXPATCH),
        ("fric_center", _D * 2), ("fric_width", _D * 2), ("outside_f", _D),
        ("face_lo", ctypes.c_int), ("nface", ctypes.c_int),
        ("f_a", _PD), ("f_b", _PD), ("f_L", _PD),
        ("f_fs", _PD), ("f_fd", _PD), ("f_dc", _PD),
        ("f_syy", _PD), ("f_tau", _PD),
        ("normal_pred", ctypes.c_int), ("fourth_order", ctypes.c_int),
        ("prim_limiting", ctypes.c_int), ("char_limiting", ctypes.c_int),
        ("dx", _D),
        ("domain", Box),
    ]


class TagParams(ctypes.Structure):
    _fields_ = [("refine_thresh", _D), ("plastic_thresh", _D), ("tag_buffer", ctypes.c_int),
                ("x_coarsen", _D * 2), ("y_coarsen", _D * 2), ("slope_coarsen", _D), ("width_coarsen", _D),
                ("x_coarsen2", _D * 2), ("y_coarsen2", _D * 2), ("level", ctypes.c_int),
                ("coarsen2_max_level", ctypes.c_int), ("stress_thresh", _D), ("sliprate_thresh", _D)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        PP = ctypes.POINTER(Params)
        PF = ctypes.POINTER(Fab)
        L.tmo_set_material.argtypes = [PP] + [_D] * 9
        for name in ("tmo_charanalysis", "tmo_charsynthesis"):
            getattr(L, name).argtypes = [PF, PP, ctypes.c_int, Box]
        L.tmo_charvalues.argtypes = [PF, PP, Box]
        L.tmo_riemann_kernel.argtypes = [PF, PF, PF, PP, ctypes.c_int, Box]
        L.tmo_getflux.argtypes = [PF, PF, PP, ctypes.c_int, Box]
        L.tmo_getadwdx.argtypes = [PF, PF, PP, _D, ctypes.c_int, Box]
        L.tmo_outbc.argtypes = [PF, PF, PP, ctypes.c_int, ctypes.c_int, Box]
        L.tmo_freebc.argtypes = [PF, PF, PP, ctypes.c_int, ctypes.c_int, Box]
        L.tmo_rsfaultbc.argtypes = [PF, PF, PF, PP, Box]
        L.tmo_swfaultbc.argtypes = [PF, PF, PF, PP, Box]
        L.tmo_rsnewton.argtypes = [PP, _D, _D, _D, _D, _PD, ctypes.POINTER(ctypes.c_int)]
        L.tmo_rsnewton.restype = _D
        L.tmo_rssetbnd.argtypes = [PF, PF, PF, PP, _D, _D, Box]
        L.tmo_swsetbnd.argtypes = [PF, PF, PF, PP, _D, _D, Box]
        L.tmo_plasticflow.argtypes = [PP, _D] + [_PD] * 8
        L.tmo_plastic_correction.argtypes = [PF, PP, _D, Box]
        L.tmo_vanleer_slopes.argtypes = [PF, PF, ctypes.c_int, ctypes.c_int, ctypes.c_int, Box, Box]
        L.tmo_ppm_face_values.argtypes = [PF, PF, ctypes.c_int, ctypes.c_int, ctypes.c_int, Box, Box]
        L.tmo_fourth_order_slopes.argtypes = [PF, PF, PF, ctypes.c_int, Box, Box]
        L.tmo_ppm_limiter.argtypes = [PF, PF, Box]
        L.tmo_ppm_normal_pred.argtypes = [PF, PF, PF, ctypes.c_double, Box]
        L.tmo_plm_normal_pred.argtypes = [PF, PF, PF, PF, ctypes.c_double, Box]
        L.tmo_update_state.argtypes = [PF, PF, PF, PF, PP, Box, _D, _D]
        L.tmo_update_state.restype = ctypes.c_int
        L.tmo_fullgrad.argtypes = [PF, PF, Box, Box]
        L.tmo_level_steps.argtypes = [_PD, _PD, PP, ctypes.c_int, ctypes.c_int, _D, _D,
                                      ctypes.c_int, ctypes.c_int]
        L.tmo_level_steps.restype = ctypes.c_int
        L.tmo_level_steps_boxed.argtypes = [_PD, _PD, PP, ctypes.c_int, ctypes.c_int, _D, _D,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tmo_level_steps_boxed.restype = ctypes.c_int
        L.tmo_tag_level.argtypes = [_PD, _PD, PP, ctypes.POINTER(TagParams), ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_ubyte)]
        L.tmo_tag_level.restype = ctypes.c_int
        _lib = L
    return _lib


class FabArray:
    """A numpy-backed FArrayBox: array shape (nc, ny, nx) in C order == Fortran (x fastest)."""

    def __init__(self, lo, hi, nc, fill=0.0, data=None):
        self.lo = tuple(lo)
        self.hi = tuple(hi)
        self.nc = nc
        shape = (nc, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        if data is None:
            self.a = np.full(shape, fill, dtype=np.float64)
        else:
            assert data.shape == shape and data.dtype == np.float64 and data.flags.c_contiguous
            self.a = data
        self._fab = Fab()
        self._fab.d = self.a.ctypes.data_as(_PD)
        self._fab.b = Box.make(lo[0], lo[1], hi[0], hi[1])
        self._fab.nc = nc

    @property
    def fab(self):
        return ctypes.byref(self._fab)

    def at(self, i, j, c):
        return self.a[c, j - self.lo[1], i - self.lo[0]]

    def view(self, lo, hi):
        """numpy view (nc, ny, nx) of the index range lo..hi (inclusive)."""
        return self.a[:, lo[1] - self.lo[1]:hi[1] - self.lo[1] + 1, lo[0] - self.lo[0]:hi[0] - self.lo[0] + 1]


def _arr(x):
    if x is None:
        return None
    x = np.ascontiguousarray(x, dtype=np.float64)
    return x


class OracleParams:
    """Owns a Params struct plus the per-face arrays it points to."""

    def __init__(self, cfg):
        self.cfg = cfg
        p = Params()
        L = lib()
        bg = cfg.get("background", (0.0,) * 6)
        L.tmo_set_material(ctypes.byref(p), cfg["cs"], cfg["cp"], cfg["G"], *bg)
        p.p_mu = cfg.get("p_mu", 0.5735)
        p.p_beta = cfg.get("p_beta", p.p_mu / 2)
        p.p_eta = cfg.get("p_eta", 0.2775)
        p.p_b = cfg.get("p_b", 0.0)
        p.p_h = cfg.get("p_h", 0.0)
        p.use_plasticity = int(cfg.get("plasticity", 0))
        for k, v in enumerate(cfg.get("bc", (0, 0, 0, 0))):
            p.bc[k] = v
        p.fault_kind = 0 if cfg.get("fault", "rs") == "rs" else 1
        rs = cfg.get("rs", {})
        for name, key, dflt in (("r0", "nuc_R", 0.0), ("nucTime", "nuc_T", 0.0), ("nsig", "nuc_S", 0.0),
                                ("rs_a", "a", 0.01), ("rs_b", "b", 0.014), ("rs_V0", "V0", 1e-6),
                                ("rs_f0", "f0", 0.6), ("rs_L", "L", 0.02), ("rs_fw", "fw", 0.13),
                                ("rs_Vw", "Vw", 0.17), ("rs_fExp", "fric_exp", -1.0), ("x0", "nuc_x", 0.0),
                                ("y0", "nuc_y", 0.0), ("rs_rupture_thr", "rupture_thr", 0.001),
                                ("ramp_x", "ramp_x", 0.0), ("ramp_w", "ramp_w", 1e40),
                                ("ramp_a", "ramp_a", 0.0), ("ramp_Vw", "ramp_Vw", 0.0),
                                ("rs_psi0", "psi", 0.6)):
            setattr(p, name, float(rs.get(key, dflt)))
        p.state_law = int(rs.get("state_law", 0))
        sw = cfg.get("sw", {})
        p.fs = sw.get("f_static", 0.677)
        p.fd = sw.get("f_dynamic", 0.525)
        p.dw = sw.get("d_weak", 0.4)
        p.smValue = sw.get("smooth", 0.0)
        p.sw_rupture_thr = sw.get("rupture_thr", 0.001)
        patches = sw.get("patches", [])
        p.num_patches = len(patches)
        for k, (xc, xw, tau) in enumerate(patches):
            p.xc[k], p.xw[k], p.tau_p[k] = xc, xw, tau
        fc = sw.get("fric_center", (0.0, 0.0))
        fw = sw.get("fric_width", (1e40, 1e40))
        p.fric_center[0], p.fric_center[1] = fc
        p.fric_width[0], p.fric_width[1] = fw
        p.outside_f = sw.get("outside_f", 1e4)
        faces = cfg.get("faces", {})
        self._keep = {}
        p.face_lo = int(faces.get("lo", 0))
        p.nface = 0
        for name in ("f_a", "f_b", "f_L", "f_fs", "f_fd", "f_dc", "f_syy", "f_tau"):
            arr = _arr(faces.get(name[2:]))
            if arr is not None:
                self._keep[name] = arr
                p.nface = len(arr)
                setattr(p, name, arr.ctypes.data_as(_PD))
        num = cfg.get("numerics", {})
        p.normal_pred = {"CTU": 0, "PLM": 1, "PPM": 2}[num.get("normal_predictor", "PPM")]
        p.fourth_order = int(num.get("fourth_order", 1))
        p.prim_limiting = int(num.get("prim_limiting", 1))
        p.char_limiting = int(num.get("char_limiting", 0))
        p.dx = cfg["dx"]
        nx, ny = cfg["nx"], cfg["ny"]
        p.domain = Box.make(0, 0, nx - 1, ny - 1)
        self.p = p

    @property
    def ref(self):
        return ctypes.byref(self.p)


def level_steps(op, U, B, time, dt, nsteps, nthreads=1, maxbox=None):
    """Advance a uniform level (arrays from level_arrays) in place; returns the error code."""
    L = lib()
    nx, ny = op.cfg["nx"], op.cfg["ny"]
    assert U.flags.c_contiguous and B.flags.c_contiguous
    if maxbox is None:
        return L.tmo_level_steps(U.ctypes.data_as(_PD), B.ctypes.data_as(_PD), op.ref, nx, ny,
                                 time, dt, nsteps, nthreads)
    return L.tmo_level_steps_boxed(U.ctypes.data_as(_PD), B.ctypes.data_as(_PD), op.ref, nx, ny,
                                   time, dt, nsteps, maxbox, nthreads)


def tag_level(op, U, B, tp_fields):
    """Tags (ny, nx) uint8 of a uniform level; tp_fields: dict of TagParams fields."""
    tp = TagParams()
    for k, v in tp_fields.items():
        if isinstance(v, (tuple, list)):
            getattr(tp, k)[:] = v
        else:
            setattr(tp, k, v)
    nx, ny = op.cfg["nx"], op.cfg["ny"]
    out = np.zeros((ny, nx), dtype=np.uint8)
    assert lib().tmo_tag_level(U.ctypes.data_as(_PD), B.ctypes.data_as(_PD), op.ref, ctypes.byref(tp), nx, ny,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))) == 0
    return out
