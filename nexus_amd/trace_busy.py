"""Kernel-busy fraction of a time window from a rocprofv3 kernel trace.

Usage: python -m nexus_amd.trace_busy KERNEL_TRACE_CSV START_NS END_NS

START_NS / END_NS are CLOCK_MONOTONIC nanoseconds bracketing the work of
interest (``bench.py --only bert`` prints them around its timed layer calls as
``window monotonic_ns A B``).  Prints the kernels inside the window, the union of
their execution intervals (the time the device ran at least one kernel), the sum
of their durations, and the gaps between kernels: the time a captured graph of
the same launches could at most recover.
"""
import csv
import sys


def busy(trace_csv, t0, t1):
    iv = []
    with open(trace_csv) as fh:
        for row in csv.DictReader(fh):
            s, e = int(row["Start_Timestamp"]), int(row["End_Timestamp"])
            if s >= t0 and e <= t1:
                iv.append((s, e))
    iv.sort()
    if not iv:
        return None
    union, gaps = 0, []
    cs, ce = iv[0]
    for s, e in iv[1:]:
        if s > ce:
            union += ce - cs
            gaps.append(s - ce)
            cs, ce = s, e
        else:
            ce = max(ce, e)
    union += ce - cs
    gaps.sort()
    return {"kernels": len(iv), "window_ms": (t1 - t0) / 1e6, "busy_ms": union / 1e6,
            "busy_frac": union / (t1 - t0), "sum_ms": sum(e - s for s, e in iv) / 1e6, "gaps": len(gaps),
            "gap_ms": sum(gaps) / 1e6, "median_gap_us": gaps[len(gaps) // 2] / 1e3 if gaps else 0.0,
            "gaps_over_20us": sum(1 for g in gaps if g > 20000),
            "gap_over_20us_ms": sum(g for g in gaps if g > 20000) / 1e6}


if __name__ == "__main__":
    if len(sys.argv) != 4:
        sys.exit(__doc__)
    r = busy(sys.argv[1], int(sys.argv[2]), int(sys.argv[3]))
    if r is None:
        sys.exit("no kernel inside the window (the trace's clock must be CLOCK_MONOTONIC)")
    for k, v in r.items():
        print("%s: %s" % (k, round(v, 4) if isinstance(v, float) else v))
