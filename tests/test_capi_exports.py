"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares."""
import ctypes
import glob
import os
import re

from tetemoko_amd import bdry_plot, capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# each header's entry points live in one library
LIBS = {"tetemoko_hip.h": (capi.LIB_PATH, capi.EXPORTED),
        "tetemoko_h5.h": (bdry_plot.LIB_PATH, bdry_plot.EXPORTED)}


def _declared(header=None):
    names = set()
    hs = [os.path.join(ROOT, "include", header)] if header else glob.glob(os.path.join(ROOT, "include", "*.h"))
    for h in hs:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(tmk_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    assert {os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))} == set(LIBS)
    names = _declared("tetemoko_hip.h")
    assert "tmk_patch_update_state" in names and "tmk_level_step" in names
    for h, (_, exported) in LIBS.items():
        names = _declared(h)
        assert names == set(exported), (h, names ^ set(exported))


def test_library_exports_all_declared():
    for h, (path, _) in LIBS.items():
        L = ctypes.CDLL(path)
        for n in _declared(h):
            assert hasattr(L, n), (h, n)


def test_strerror_without_device():
    assert capi.strerror(capi.E_PSI_NAN).startswith("NaN")
    assert capi.strerror(-99) == "unknown error"
