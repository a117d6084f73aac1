"""The kernel-trace summary's busy time: the union of overlapping kernel intervals."""
import os

from tetemoko_amd.trace_summary import summarize


def test_union_of_overlapping_kernels(tmp_path):
    d = tmp_path / "run"
    os.makedirs(d)
    rows = [("k_a(int)", 0, 100), ("k_b(int)", 50, 150), ("k_a(int)", 400, 500), ("k_c", 1000, 1000000)]
    with open(d / "x_kernel_trace.csv", "w") as fh:
        fh.write("Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for n, s, e in rows:
            fh.write('"%s",%d,%d\n' % (n, s, e))
    out = summarize(str(tmp_path))
    # busy: [0, 150] + [400, 500] + [1000, 1000000] = 150 + 100 + 999000 ns of a 1e6 ns span
    assert "span 0.00 s" in out.splitlines()[0]
    assert "(99.9 %)" in out.splitlines()[0]
    assert "idle gaps >= 10 us: 0" in out
    assert "k_a" in out and "k_b" in out
