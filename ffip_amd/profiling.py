"""Helpers for rocprofv3 measurements of the engine.

    python -m ffip_amd.profiling copy [GiB]        # calibration workload: a plain device copy
    python -m ffip_amd.profiling summarize DIR     # per-kernel table of a rocprofv3 CSV output

`copy` copies a buffer of known size a few times with torch (one copy kernel
per call), so that the FETCH_SIZE / WRITE_SIZE counters of that kernel can be
compared with the bytes it must move: the ratio is the scale factor applied to
the same counters read on the FDTD kernels.
"""
from __future__ import annotations

import csv
import glob
import os
import sys
from collections import defaultdict


def copy(gib: float = 1.0, reps: int = 5):
    import torch
    n = int(gib * (1 << 30)) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied %d x %.3f GiB (%d bytes read and written per copy)" % (reps, gib, n * 4))


def summarize(d: str):
    """Print mean duration (kernel_stats) and mean counter values per kernel."""
    out = []
    stats = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    for path in stats:
        with open(path) as f:
            for r in csv.DictReader(f):
                out.append("| %s | %s | %.2f | %.2f |" % (r["Name"][:70], r["Calls"], float(r["AverageNs"]) / 1000,
                                                        float(r["Percentage"])))
    if out:
        print("| kernel | calls | average us | % of GPU time |")
        print("|---|---|---|---|")
        print("\n".join(out))
    counters = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    acc = defaultdict(list)
    for path in counters:
        with open(path) as f:
            for r in csv.DictReader(f):
                acc[(r["Kernel_Name"][:70], r["Counter_Name"])].append(float(r["Counter_Value"]))
    if acc:
        print()
        print("| kernel | counter | dispatches | mean per dispatch |")
        print("|---|---|---|---|")
        for (k, c), v in sorted(acc.items()):
            print("| %s | %s | %d | %.1f |" % (k, c, len(v), sum(v) / len(v)))


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "copy":
        copy(float(sys.argv[2]) if len(sys.argv) > 2 else 1.0)
    elif len(sys.argv) >= 3 and sys.argv[1] == "summarize":
        summarize(sys.argv[2])
    else:
        print(__doc__)
        sys.exit(2)
