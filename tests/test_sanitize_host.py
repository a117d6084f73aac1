"""The library's host code under AddressSanitizer, UndefinedBehaviorSanitizer and LeakSanitizer.

``tests/sanitize`` builds ``tmk_api.hip`` and ``tmk_f32.hip`` for the host only, with the
sanitizers on, links them against a host-memory stand-in of the HIP runtime calls they make and
drives every list builder (boxes, islands, tiles, strips, exchange rectangles, coarse-fine fill,
flux registers, average-down, regrid, batch and hierarchy membership) through the C ABI on
random box layouts.  No GPU code runs: the check is that the host code stays inside its
allocations, has no undefined behaviour, frees every device allocation it makes and returns
TMK_OK where the inputs are valid.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def _toolchain():
    return os.path.exists("/opt/rocm/bin/hipcc") and os.path.exists("/opt/rocm/llvm/bin/clang++") and shutil.which("make")


@pytest.mark.skipif(not _toolchain(), reason="needs ROCm's hipcc and clang++ and make")
def test_list_builders_clean_under_sanitizers():
    b = subprocess.run(["make", "-j4"], cwd=HERE, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout[-4000:] + b.stderr[-4000:]
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([os.path.join(HERE, "_build", "list_builders"), "12"], cwd=HERE, capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-8000:]
    assert "rounds clean" in r.stdout
