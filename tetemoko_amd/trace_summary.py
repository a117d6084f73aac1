"""Device-busy summary of a rocprofv3 kernel trace.

    rocprofv3 --kernel-trace -d DIR -o run --output-format csv -- python3 -m tetemoko_amd.serving ...
    python3 -m tetemoko_amd.trace_summary DIR

Reports the span from the first kernel's start to the last kernel's end, the time at least one
kernel was running (the union of the kernel intervals, so concurrent launches on several streams
count once), the idle gaps between them, and the total kernel time per kernel name.
"""
import collections
import csv
import glob
import os
import sys


def summarize(trace_dir, top=20):
    files = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise FileNotFoundError("no *kernel_trace.csv under %s" % trace_dir)
    iv, per, cnt = [], collections.Counter(), collections.Counter()
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            iv.append((s, e))
            name = r["Kernel_Name"].split("(")[0]
            per[name] += e - s
            cnt[name] += 1
    iv.sort()
    busy, gaps = 0, []
    cs, ce = iv[0]
    for s, e in iv[1:]:
        if s > ce:
            busy += ce - cs
            gaps.append(s - ce)
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    span = iv[-1][1] - iv[0][0]
    lines = ["kernels %d, span %.2f s, device busy %.2f s (%.1f %%)" % (len(iv), span / 1e9, busy / 1e9,
                                                                      100.0 * busy / span)]
    for th in (1e4, 1e5, 1e6):
        big = [g for g in gaps if g >= th]
        lines.append("idle gaps >= %g us: %d, %.2f s" % (th / 1e3, len(big), sum(big) / 1e9))
    tot = sum(per.values())
    lines.append("kernel time %.2f s (concurrent launches overlap)" % (tot / 1e9))
    for n, t in per.most_common(top):
        lines.append("  %8.3f s %5.1f %% %8d  %s" % (t / 1e9, 100.0 * t / tot, cnt[n], n))
    return "\n".join(lines)


if __name__ == "__main__":
    print(summarize(sys.argv[1]))
