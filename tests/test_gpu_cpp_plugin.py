"""The C++ plugin-surface mirror (tetemoko_amd/host/LinElastHIP.H) driven like the
reference's LinElastLevelGodunov, checked against the oracle (tests/cpp/test_plugin.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_plugin")


@pytest.mark.gpu
def test_cpp_plugin_classes():
    assert os.path.exists(BIN), "build first: __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
