"""The C-ABI library loads without a GPU and exports every entry point include/*.h declares."""
import ctypes
import glob
import os
import re

from tetemoko_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(tmk_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "tmk_patch_update_state" in names and "tmk_level_step" in names
    assert names == set(capi.EXPORTED), names ^ set(capi.EXPORTED)


def test_library_exports_all_declared():
    L = ctypes.CDLL(capi.LIB_PATH)
    for n in _declared():
        assert hasattr(L, n), n


def test_strerror_without_device():
    assert capi.strerror(capi.E_PSI_NAN).startswith("NaN")
    assert capi.strerror(-99) == "unknown error"
