"""Boundary plot files: the fault data of every level in Chombo's AMR HDF5 layout.

AMRLevelLinElast::writeCustomPlotFile writes, every plot_interval coarse steps and at the end
of a run, "<prefix>boundary.<step %06d>.2d.hdf5" with a root header (max_level, num_levels,
iteration, time, num_components, component_<n>) and, for each level that has fault boxes, a
group level_<l> (ref_ratio, dx, dt, time, prob_domain = the fault face box of the level's
domain) holding that level's fault data over its fault boxes (AMRLevelLinElast.cpp:1628-1756).
The file is written by libtmk_h5.so (host C++ over the HDF5 C library); the fault data come
from the device through tmk_level_download_fault.
"""
import ctypes
import os

import numpy as np

from . import scenarios
from .stations import BDRY_NAMES

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libtmk_h5.so")
G = scenarios.GHOST

_PD = ctypes.POINTER(ctypes.c_double)


class BdryLevel(ctypes.Structure):
    """tmk_bdry_level (include/tetemoko_h5.h)."""
    _fields_ = [("level", ctypes.c_int), ("ref_ratio", ctypes.c_int), ("dx", ctypes.c_double),
                ("dt", ctypes.c_double), ("time", ctypes.c_double), ("domain_nx", ctypes.c_int),
                ("nbox", ctypes.c_int), ("boxes", ctypes.POINTER(ctypes.c_int)), ("data", ctypes.POINTER(_PD))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtmk_h5.so is not built (tetemoko_amd/_lib); run __graft_entry__.build() "
                               "with the HDF5 headers present")
        L = ctypes.CDLL(LIB_PATH)
        L.tmk_h5_write_bdry_plot.restype = ctypes.c_int
        L.tmk_h5_write_bdry_plot.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                             ctypes.c_int, ctypes.POINTER(BdryLevel)]
        L.tmk_h5_bdry_plot_name.restype = ctypes.c_int
        L.tmk_h5_bdry_plot_name.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        _lib = L
    return _lib


EXPORTED = ("tmk_h5_write_bdry_plot", "tmk_h5_bdry_plot_name")


class H5Error(RuntimeError):
    pass


def plot_name(prefix, step):
    buf = ctypes.create_string_buffer(4096)
    n = lib().tmk_h5_bdry_plot_name(prefix.encode(), int(step), buf, len(buf))
    if n < 0:
        raise H5Error("plot file name too long")
    return buf.value.decode()


def write(path, levels, step, time, names, max_level=None):
    """Write one boundary plot file.  levels: dicts with level, ref_ratio, dx, dt, time,
    domain_nx, boxes [(x0, 0, x1, 0)] (the fault boxes) and data [per box an array
    (ncomp, x1 - x0 + 1 + 8), the box's fault data with its 4 x-ghosts]."""
    ncomp = len(names)
    keep = []
    arr = (BdryLevel * max(1, len(levels)))()
    for k, lv in enumerate(levels):
        boxes = np.ascontiguousarray(np.asarray(lv["boxes"], dtype=np.int32).reshape(-1, 4))
        datas = [np.ascontiguousarray(d, dtype=np.float64) for d in lv["data"]]
        if len(datas) != boxes.shape[0]:
            raise H5Error("one data array per fault box")
        for b, d in zip(boxes, datas):
            if d.shape != (ncomp, b[2] - b[0] + 1 + 2 * G):
                raise H5Error("fault data of box %s has shape %s" % (tuple(b), d.shape))
        ptrs = (_PD * max(1, len(datas)))(*[d.ctypes.data_as(_PD) for d in datas])
        keep += [boxes, datas, ptrs]
        arr[k] = BdryLevel(int(lv["level"]), int(lv["ref_ratio"]), float(lv["dx"]), float(lv["dt"]),
                           float(lv["time"]), int(lv["domain_nx"]), boxes.shape[0],
                           boxes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ptrs)
    finest = len(levels) - 1
    cnames = (ctypes.c_char_p * ncomp)(*[n.encode() for n in names])
    code = lib().tmk_h5_write_bdry_plot(path.encode(), finest if max_level is None else int(max_level),
                                        finest, int(step), float(time), ncomp, cnames, len(levels), arr)
    if code != 0:
        raise H5Error("tmk_h5_write_bdry_plot(%s): error %d" % (path, code))
    return path


def level_fault_data(level, boxes, nb):
    """The fault boxes of a device level (boxes with y0 == 0, as face boxes (x0, 0, x1, 0)) and
    their fault data downloaded from the device."""
    fboxes, data = [], []
    for k, b in enumerate(boxes):
        if b[1] != 0:
            continue
        d = np.empty((nb, b[2] - b[0] + 1 + 2 * G))
        level.download_fault(k, d)
        fboxes.append((b[0], 0, b[2], 0))
        data.append(d)
    return fboxes, data


def write_level(prefix, level, cfg, step, time, dt, max_level=0):
    """Boundary plot file of one device level (a single-level run)."""
    nb = scenarios.NB[cfg["fault"]]
    fb, data = level_fault_data(level, level.boxes, nb)
    lv = {"level": 0, "ref_ratio": 1, "dx": cfg["dx"], "dt": dt, "time": time, "domain_nx": cfg["nx"],
          "boxes": fb, "data": data}
    return write(plot_name(prefix, step), [lv], step, time, BDRY_NAMES[cfg["fault"]], max_level)


def write_hierarchy(prefix, H, max_level=None):
    """Boundary plot file of an amr.Hierarchy after its last coarse step (every level at the
    hierarchy's time; level l steps dt0 / r^l)."""
    H.sync()
    levels = []
    for l in range(len(H.levels)):
        cfg = H.cfgs[l]
        nb = scenarios.NB[cfg["fault"]]
        fb, data = level_fault_data(H.levels[l], H.boxes[l], nb)
        levels.append({"level": l, "ref_ratio": H.r, "dx": cfg["dx"], "dt": H.dt0 / H.r ** l, "time": H.time,
                       "domain_nx": cfg["nx"], "boxes": fb, "data": data})
    max_level = len(H.cfgs) - 1 if max_level is None else max_level
    return write(plot_name(prefix, H.steps), levels, H.steps, H.time, BDRY_NAMES[H.cfgs[0]["fault"]], max_level)
