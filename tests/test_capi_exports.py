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


def _cls(n):
    """Bytes the level-array cache allocates for a request of n: n rounded up to a multiple of
    1/8 of the largest power of two not above it (at least 256 bytes)."""
    if n < 256:
        return 256
    g = max(256, (1 << (n.bit_length() - 1)) >> 3)
    return -(-n // g) * g


def test_size_classes_waste_under_an_eighth():
    for n in [1, 255, 256, 257, 4095, 4096, 4097, 12345, 1 << 20, (1 << 20) + 1, 3 * (1 << 30) + 7]:
        c = _cls(n)
        assert c >= n and (n < 2048 or c - n < n / 8)


def test_footprint_without_a_context():
    """tmk_level_footprint_params is host code: the device bytes of a level before any context
    (each array as the level-array cache allocates it)."""
    from tetemoko_amd import scenarios
    cfg = scenarios.config_c1(seed=1, nx=96, ny=48, dx=25.0)
    g = (96 + 8) * (48 + 8)
    fault = 2 * _cls((96 + 8) * 11 * 8) + _cls((96 + 8) * 4)
    assert capi.level_footprint_params(cfg, [(0, 0, 95, 47)]) == 3 * _cls(22 * g * 8) + _cls(g * 4) + fault
    assert (capi.level_footprint_params(cfg, [(0, 0, 95, 47)], capi.PREC_FP32_BULK) ==
            3 * _cls(22 * g * 4) + _cls(g * 4) + fault)
    # two boxes far apart are two islands, each over its own ghosted box; only one has fault data
    a, b = (16 + 8) * (16 + 8), (32 + 8) * (24 + 8)
    want = (3 * (_cls(22 * a * 8) + _cls(22 * b * 8)) + _cls(a * 4) + _cls(b * 4) +
            2 * _cls((16 + 8) * 11 * 8) + _cls((16 + 8) * 4))
    assert capi.level_footprint_params(cfg, [(0, 0, 15, 15), (64, 24, 95, 47)]) == want
