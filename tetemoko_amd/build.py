"""Build libtetemoko_hip.so for gfx950 in-tree (tetemoko_amd/_lib)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument"]


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building " + what)


def build(verbose=True):
    """tmk_api.hip (host code and every FP64 kernel, no FMA contraction) and tmk_f32.hip (the
    FP32-bulk update kernels, FMA contraction on), linked into one shared library."""
    csrc = os.path.join(HERE, "csrc")
    out_dir = os.path.join(HERE, "_lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libtetemoko_hip.so")
    objs = []
    for name, contract in (("tmk_api", "off"), ("tmk_f32", "fast")):
        obj = os.path.join(out_dir, name + ".o")
        flags = [f if f != "-ffp-contract=off" else "-ffp-contract=" + contract for f in FLAGS if f != "-shared"]
        _run([HIPCC] + flags + ["-c", "-o", obj, os.path.join(csrc, name + ".hip")], verbose, name + ".hip")
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose, "libtetemoko_hip.so")
    os.replace(tmp, out)
    build_copy_calib(verbose)
    build_h5(verbose)
    return out


HDF5_ROOT = os.environ.get("TMK_HDF5_ROOT", "/opt/conda")


def build_h5(verbose=True):
    """libtmk_h5.so: the boundary plot-file writer (host C++ against the HDF5 C library of
    HDF5_ROOT, found at run time through the rpath).  Skipped, with a message, when the HDF5
    headers are absent; tetemoko_amd.bdry_plot then raises when it is used."""
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print("libtmk_h5.so skipped: no hdf5.h under " + inc, flush=True)
        return None
    out = os.path.join(HERE, "_lib", "libtmk_h5.so")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + inc, "-o", out + ".tmp", os.path.join(HERE, "csrc", "tmk_h5.cpp"),
           "-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir]
    _run(cmd, verbose, "tmk_h5.cpp")
    os.replace(out + ".tmp", out)
    return out


def build_copy_calib(verbose=True):
    """The copy kernel of known size used to scale HBM byte counters (tetemoko_amd/_lib/copy_calib)."""
    src = os.path.join(HERE, "csrc", "copy_calib.hip")
    out = os.path.join(HERE, "_lib", "copy_calib")
    flags = [f for f in FLAGS if f not in ("-fPIC", "-shared")]
    cmd = [HIPCC] + flags + ["-o", out + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building copy_calib")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
