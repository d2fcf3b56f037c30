"""Helpers around rocprofv3 output for the roofline figures of bench.py.

* ``copy_probe(nbytes)`` runs plain device-to-device copies of a known size,
  the calibration for the FETCH_SIZE / WRITE_SIZE counters.
* ``kernel_stats(dir)`` reads ``*_kernel_stats.csv`` of a
  ``rocprofv3 --kernel-trace --stats --output-format csv`` run.
* ``counter_means(dir, pattern)`` averages per-dispatch counter values of
  ``rocprofv3 --pmc ... --output-format csv`` over the matching kernels.
* ``window_*`` restrict either to the dispatches between the first and the
  last dispatch of a marker kernel (bench.py computes modularity with
  ``k_intra`` right after the load and after the last step, so the window is
  the warmup and timed batches without the initial load).
* ``summary_text`` formats them as the short text kept beside the results.
"""
from __future__ import annotations

import csv
import glob
import os
import re
from collections import defaultdict


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    return hits[0] if hits else None


def short_name(name: str) -> str:
    n = name.replace("hl::(anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    n = re.sub(r"rocprim::ROCPRIM_\d+_NS::detail::", "rocprim::", n)
    return n


def kernel_stats(d):
    path = _find(d, "kernel_stats.csv")
    if not path:
        return []
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append({"name": short_name(r["Name"]), "calls": int(r["Calls"]),
                         "total_ns": float(r["TotalDurationNs"]), "avg_ns": float(r["AverageNs"]),
                         "pct": float(r["Percentage"])})
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def kernel_avg_us(d, pattern):
    """Calls and call-weighted average duration of every kernel matching pattern."""
    rx = re.compile(pattern)
    calls, tot = 0, 0.0
    for r in kernel_stats(d):
        if rx.search(r["name"]):
            calls += r["calls"]
            tot += r["total_ns"]
    return calls, (tot / calls / 1000.0 if calls else 0.0)


def counter_means(d, pattern):
    """{counter: (dispatches, mean value)} over dispatches of kernels matching pattern."""
    path = _find(d, "counter_collection.csv")
    if not path:
        return {}
    rx = re.compile(pattern)
    acc = defaultdict(lambda: [0, 0.0])
    with open(path) as f:
        for r in csv.DictReader(f):
            if not rx.search(short_name(r.get("Kernel_Name", ""))):
                continue
            a = acc[r["Counter_Name"]]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    return {k: (v[0], v[1] / v[0]) for k, v in acc.items() if v[0]}


def _window(rows, marker, key):
    rows.sort(key=lambda r: int(r[key]))
    idx = [i for i, r in enumerate(rows) if marker in r["Kernel_Name"]]
    if len(idx) < 2:
        return []
    return rows[idx[0] + 1:idx[-1]]


def window_kernel_stats(d, marker="k_intra"):
    """[(name, calls, avg_us, total_ms)] of the kernels between the marker dispatches."""
    path = _find(d, "kernel_trace.csv")
    if not path:
        return [], 0.0
    with open(path) as f:
        rows = _window(list(csv.DictReader(f)), marker, "Start_Timestamp")
    if not rows:
        return [], 0.0
    span = (int(rows[-1]["End_Timestamp"]) - int(rows[0]["Start_Timestamp"])) / 1e6
    acc = defaultdict(lambda: [0, 0])
    for r in rows:
        a = acc[short_name(r["Kernel_Name"])]
        a[0] += 1
        a[1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    out = [(k, v[0], v[1] / v[0] / 1000.0, v[1] / 1e6) for k, v in acc.items()]
    out.sort(key=lambda t: -t[3])
    return out, span


def window_counter_means(d, pattern, marker="k_intra"):
    """counter_means restricted to the dispatches between the marker dispatches
    (the marker must be collected too, e.g. by the include regex)."""
    path = _find(d, "counter_collection.csv")
    if not path:
        return {}
    with open(path) as f:
        rows = _window(list(csv.DictReader(f)), marker, "Start_Timestamp")
    rx = re.compile(pattern)
    acc = defaultdict(lambda: [0, 0.0])
    for r in rows:
        if rx.search(short_name(r["Kernel_Name"])):
            a = acc[r["Counter_Name"]]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    return {k: (v[0], v[1] / v[0]) for k, v in acc.items() if v[0]}


def copy_probe(nbytes: int, reps: int = 4) -> None:
    """hipMemcpy device-to-device of nbytes, reps times (a plain copy kernel)."""
    from . import hip
    import numpy as np
    n = nbytes // 8
    a = hip.DeviceArray(np.ones(n, dtype=np.float64))
    b = hip.DeviceArray(np.zeros(n, dtype=np.float64))
    for _ in range(reps):
        hip.copy_d2d(b, a, n * 8)
    hip.synchronize()
    a.free()
    b.free()


def summary_text(d, command, top=20, pattern=None):
    lines = [f"command: {command}", ""]
    rows = kernel_stats(d)
    if rows:
        lines.append(f"{'kernel':80s} {'calls':>7s} {'avg_us':>9s} {'total_ms':>9s} {'pct':>6s}")
        for r in rows[:top]:
            lines.append(f"{r['name'][:80]:80s} {r['calls']:7d} {r['avg_ns'] / 1000:9.1f} "
                         f"{r['total_ns'] / 1e6:9.2f} {r['pct']:6.2f}")
        if pattern:
            c, avg = kernel_avg_us(d, pattern)
            lines += ["", f"kernels matching /{pattern}/: {c} calls, call-weighted average {avg:.2f} us"]
    win, span = window_kernel_stats(d)
    if win:
        lines += ["", f"between the modularity passes after the load and after the last step "
                      f"({span:.2f} ms of GPU time line):",
                  f"{'kernel':80s} {'calls':>7s} {'avg_us':>9s} {'total_ms':>9s}"]
        for name, calls, avg, tot in win[:top]:
            lines.append(f"{name[:80]:80s} {calls:7d} {avg:9.1f} {tot:9.2f}")
        if pattern:
            rx = re.compile(pattern)
            c = sum(w[1] for w in win if rx.search(w[0]))
            t = sum(w[3] for w in win if rx.search(w[0]))
            if c:
                lines.append(f"kernels matching /{pattern}/ in this window: {c} calls, "
                             f"call-weighted average {t * 1000.0 / c:.2f} us")
    if pattern:
        cm = counter_means(d, pattern)
        for k, (nd, mean) in sorted(cm.items()):
            lines.append(f"counter {k} (KiB) over all {nd} dispatches of /{pattern}/: mean {mean:.1f}")
        cw = window_counter_means(d, pattern)
        for k, (nd, mean) in sorted(cw.items()):
            lines.append(f"counter {k} (KiB) over {nd} dispatches of /{pattern}/ after the load: mean {mean:.1f}")
    return "\n".join(lines) + "\n"


def traffic_entry(fetch_dir, write_dir, pattern, fetch_factor, write_factor, kernel):
    """HBM bytes per launch of the kernels matching pattern, from two --pmc
    runs (FETCH_SIZE, WRITE_SIZE), dispatches after the load only."""
    fm = window_counter_means(fetch_dir, pattern).get("FETCH_SIZE")
    wm = window_counter_means(write_dir, pattern).get("WRITE_SIZE")
    if not fm or not wm:
        return None
    return {"kernel": kernel, "dispatches": fm[0], "fetch_kib": round(fm[1], 1), "write_kib": round(wm[1], 1),
            "bytes_per_launch": int(round(fm[1] * 1024 * fetch_factor + wm[1] * 1024 * write_factor))}


def calibration(copy_fetch_dir, copy_write_dir, copy_bytes):
    """True/counted byte ratios of a device-to-device copy of copy_bytes."""
    f = counter_means(copy_fetch_dir, "copyBuffer").get("FETCH_SIZE")
    w = counter_means(copy_write_dir, "copyBuffer").get("WRITE_SIZE")
    ff = copy_bytes / (f[1] * 1024) if f and f[1] > 0 else 1.0
    wf = copy_bytes / (w[1] * 1024) if w and w[1] > 0 else 1.0
    return {"copy_bytes": copy_bytes, "fetch_kib": round(f[1], 1) if f else None,
            "write_kib": round(w[1], 1) if w else None, "fetch_factor": round(ff, 4), "write_factor": round(wf, 4)}


def main(argv=None):
    """python -m hit_leiden_amd.profiling summary DIR --command C [--kernel RX]
    python -m hit_leiden_amd.profiling traffic --out F --fetch D --write D
        --copy-fetch D --copy-write D --copy-bytes N --command C SLOT=REGEX ..."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="hit_leiden_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("summary")
    a.add_argument("dir")
    a.add_argument("--command", default="")
    a.add_argument("--kernel", default=None)
    b = sub.add_parser("traffic")
    b.add_argument("--out", required=True)
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--copy-fetch", required=True)
    b.add_argument("--copy-write", required=True)
    b.add_argument("--copy-bytes", type=int, required=True)
    b.add_argument("--command", default="")
    b.add_argument("slots", nargs="+", help="slot=kernel regex")
    args = ap.parse_args(argv)
    if args.cmd == "summary":
        print(summary_text(args.dir, args.command, pattern=args.kernel), end="")
        return
    cal = calibration(args.copy_fetch, args.copy_write, args.copy_bytes)
    out = {"_method": "per-dispatch FETCH_SIZE and WRITE_SIZE (KiB) from two rocprofv3 --pmc runs of '"
                      + args.command + "', averaged over the dispatches after the initial load, each scaled by "
                      "the ratio of true bytes to the counter for a device-to-device copy of "
                      f"{args.copy_bytes} bytes", "_calibration": cal}
    for s in args.slots:
        slot, rx = s.split("=", 1)
        e = traffic_entry(args.fetch, args.write, rx, cal["fetch_factor"], cal["write_factor"], rx)
        if e:
            out[slot] = e
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
