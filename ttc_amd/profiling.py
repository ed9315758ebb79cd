"""Summaries of rocprofv3 output for the transposition kernels.

  python -m ttc_amd.profiling stats DIR [-o OUT.md]
      per-kernel launches, total and average duration from a
      `rocprofv3 --kernel-trace --stats` output directory.
      With --kernel K --last N, also the durations of the last N dispatches
      of kernel K from *kernel_trace.csv (the region that `bench.py` times
      for its roofline line comes last among that kernel's launches).
  python -m ttc_amd.profiling counters DIR [-o OUT.md]
      per-kernel average of every counter of a `rocprofv3 --pmc` run
      (for example SQ_LDS_BANK_CONFLICT against SQ_ACTIVE_INST_LDS).
  python -m ttc_amd.profiling traffic --fetch DIR --write DIR
         --calib-fetch DIR --calib-write DIR --calib-bytes N [--last N]
         [--kernel K --kernel-last N] [-o OUT]
      HBM bytes per launch of every kernel from two `rocprofv3 --pmc` runs
      (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what
      the same counter reads for a plain copy of N bytes (`bench.py
      --calibrate-copy`).  Each kernel's last --last dispatches, kernel K's
      last --kernel-last (its roofline window).  Writes OUT.md and OUT.json.

The counters read kilobytes of L2-to-memory traffic per dispatch; on gfx950
FETCH_SIZE counts 128-byte read requests as 64 bytes, which the copy
calibration corrects for streams of full lines (the factor is printed).
"""
import argparse
import csv
import glob
import json
import os
import re
import sys


def short_name(kernel):
    """'void ttc::transpose_rc_kernel<ttc::OpDD, 1, 1, true, false>(ttc::RCParams, ...)'
    -> 'ttc::transpose_rc_kernel<ttc::OpDD, 1, 1, true, false>'."""
    k = kernel.strip()
    if k.startswith("void "):
        k = k[5:]
    depth = 0
    for i, ch in enumerate(k):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return k[:i]
    return k


def _one(directory, pattern):
    files = sorted(glob.glob(os.path.join(directory, "**", pattern), recursive=True))
    if not files:
        raise FileNotFoundError("no %s under %s" % (pattern, directory))
    return files[0]


def kernel_stats(directory):
    """[(name, calls, total_ns, avg_ns)] from *kernel_stats.csv, by total time."""
    rows = []
    with open(_one(directory, "*kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"])))
    rows.sort(key=lambda x: -x[2])
    return rows


def last_dispatches(directory, kernel, n):
    """(avg duration ns, min, max, span ns / n) of the last n dispatches of
    `kernel` in *kernel_trace.csv; span runs from the first start to the
    last end, so span / n includes the gaps between launches."""
    rows = []
    with open(_one(directory, "*kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            if short_name(r["Kernel_Name"]) == kernel:
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    rows.sort()
    rows = rows[-n:]
    if not rows:
        raise ValueError("no dispatch of %s" % kernel)
    d = [e - b for b, e in rows]
    return sum(d) / len(d), min(d), max(d), (rows[-1][1] - rows[0][0]) / len(rows), len(rows)


def counter_per_kernel(directory, counter):
    """{kernel: [value per dispatch]} of one counter from *counter_collection.csv."""
    rows = []
    with open(_one(directory, "*counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                rows.append((int(r["Dispatch_Id"]), short_name(r["Kernel_Name"]), float(r["Counter_Value"])))
    rows.sort()  # dispatch order
    out = {}
    for _, k, v in rows:
        out.setdefault(k, []).append(v)
    return out


def _copy_kernel(per_kernel):
    names = [k for k in per_kernel if re.search("copy", k, re.I) and not re.search("fill|distribution", k, re.I)]
    if not names:
        raise ValueError("no copy kernel in the calibration run")
    return max(names, key=lambda k: len(per_kernel[k]))


def traffic(fetch_dir, write_dir, calib_fetch_dir, calib_write_dir, calib_bytes, last=0, kernel=None, kernel_last=0):
    """Bytes per launch of each kernel, from counter runs scaled by the copy
    calibration: each kernel's last `last` dispatches (0: all), and for
    `kernel` its last `kernel_last` (the window its roofline line times)."""
    cf = counter_per_kernel(calib_fetch_dir, "FETCH_SIZE")
    cw = counter_per_kernel(calib_write_dir, "WRITE_SIZE")
    kf, kw = _copy_kernel(cf), _copy_kernel(cw)
    fetch_kib = sorted(cf[kf])[len(cf[kf]) // 2]
    write_kib = sorted(cw[kw])[len(cw[kw]) // 2]
    f_scale = calib_bytes / (fetch_kib * 1024.0)
    w_scale = calib_bytes / (write_kib * 1024.0)
    fetch = counter_per_kernel(fetch_dir, "FETCH_SIZE")
    write = counter_per_kernel(write_dir, "WRITE_SIZE")
    def window(k, v):
        n = kernel_last if (kernel and kernel_last and k == kernel) else last
        return v[-n:] if n else v
    fetch = {k: window(k, v) for k, v in fetch.items()}
    write = {k: window(k, v) for k, v in write.items()}
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        if "ttc::" not in k:
            continue
        f = sum(fetch[k]) / len(fetch[k]) * 1024.0 * f_scale
        w = sum(write[k]) / len(write[k]) * 1024.0 * w_scale
        kernels[k] = {"launches": len(fetch[k]), "fetch_bytes": round(f), "write_bytes": round(w),
                      "traffic_bytes": round(f + w)}
    return {"calibration": {"copy_bytes": calib_bytes, "fetch_kib": fetch_kib, "write_kib": write_kib,
                            "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
                            "copy_kernel": kf},
            "kernels": kernels}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("-o", "--out")
    s.add_argument("--command", default="", help="the command that produced DIR (recorded in the summary)")
    s.add_argument("--kernel", help="kernel whose last --last dispatches are summarised")
    s.add_argument("--last", type=int, default=0)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("-o", "--out")
    c.add_argument("--command", default="", help="the command that produced DIR (recorded in the summary)")
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--calib-fetch", "--calib-write"):
        t.add_argument(a, required=True)
    t.add_argument("--calib-bytes", type=int, required=True)
    t.add_argument("--last", type=int, default=0, help="use only each kernel's last N dispatches")
    t.add_argument("--kernel", help="kernel summarised over its own window (--kernel-last)")
    t.add_argument("--kernel-last", type=int, default=0, help="that kernel's last N dispatches")
    t.add_argument("-o", "--out", required=True, help="output path without extension")
    t.add_argument("--command", default="", help="the commands that produced the directories")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        rows = kernel_stats(args.dir)
        total = sum(r[2] for r in rows) or 1.0
        lines = ["# rocprofv3 kernel statistics", "", "Command: `%s`" % args.command if args.command else "", "",
                 "| kernel | launches | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
        for name, calls, tot, avg in rows:
            lines.append("| `%s` | %d | %.3f | %.2f | %.1f |" % (name, calls, tot / 1e6, avg / 1e3, 100 * tot / total))
        if args.kernel and args.last:
            avg, lo, hi, span, n = last_dispatches(args.dir, args.kernel, args.last)
            lines += ["", "Last %d dispatches of `%s`: average duration %.2f us (min %.2f, max %.2f); "
                      "first start to last end / %d = %.2f us per launch, gaps included." % (
                          n, args.kernel, avg / 1e3, lo / 1e3, hi / 1e3, n, span / 1e3)]
        text = "\n".join(lines) + "\n"
        if args.out:
            with open(args.out, "w") as f:
                f.write(text)
        sys.stdout.write(text)
        return 0
    if args.cmd == "counters":
        per = {}
        with open(_one(args.dir, "*counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                k = short_name(r["Kernel_Name"])
                if "ttc::" not in k:
                    continue
                per.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        names = sorted({n for v in per.values() for n in v})
        lines = ["# rocprofv3 counters per kernel (average per dispatch)", ""]
        if args.command:
            lines += ["Command: `%s`" % args.command, ""]
        lines += ["| kernel | dispatches | " + " | ".join(names) + " |", "|---|---:|" + "---:|" * len(names)]
        for k, v in sorted(per.items()):
            n = len(next(iter(v.values())))
            lines.append("| `%s` | %d | " % (k, n) + " | ".join(
                "%.4g" % (sum(v[x]) / len(v[x])) if x in v else "" for x in names) + " |")
        text = "\n".join(lines) + "\n"
        if args.out:
            with open(args.out, "w") as f:
                f.write(text)
        sys.stdout.write(text)
        return 0
    res = traffic(args.fetch, args.write, args.calib_fetch, args.calib_write, args.calib_bytes, args.last,
                  args.kernel, args.kernel_last)
    res["last_dispatches"] = args.last or None
    if args.kernel and args.kernel_last:
        res["kernel_window"] = {"kernel": args.kernel, "last_dispatches": args.kernel_last}
    if args.command:
        res["command"] = args.command
    with open(args.out + ".json", "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    c = res["calibration"]
    lines = ["# HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", ""]
    if args.command:
        lines += ["Commands: `%s`" % args.command, ""]
    lines += ["Calibration: a plain copy of %d bytes (`%s`) reads FETCH_SIZE = %.1f KiB and WRITE_SIZE = %.1f KiB;"
              " scale factors %.4f (fetch) and %.4f (write)." % (c["copy_bytes"], c["copy_kernel"], c["fetch_kib"],
                                                                 c["write_kib"], c["fetch_scale"], c["write_scale"]),
              "", (("Each kernel's last %d dispatches." % args.last) if args.last else "All dispatches.") + (
                  (" `%s`: its last %d dispatches (the launches its roofline line times)." % (
                      args.kernel, args.kernel_last)) if args.kernel and args.kernel_last else ""),
              "", "| kernel | launches | fetch MB/launch | write MB/launch | traffic MB/launch |",
              "|---|---:|---:|---:|---:|"]
    for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["traffic_bytes"] * kv[1]["launches"]):
        lines.append("| `%s` | %d | %.2f | %.2f | %.2f |" % (k, v["launches"], v["fetch_bytes"] / 1e6,
                                                            v["write_bytes"] / 1e6, v["traffic_bytes"] / 1e6))
    with open(args.out + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main())
