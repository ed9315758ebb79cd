"""Summaries of the tuning database (TTC's ttc.db with the GPU columns).

The reference analyses its database through the 'fastest0Percent' view
(ttc/ttc.py:205-225) and plots (generatePlots.py:141-244).  This module
answers the same questions as tables: per type pair and rank, how the best
bandwidth found is distributed; which kernel family, vector widths and tile
numbering won; how often the planner's own choice was already the fastest,
and what the search gained over it.

  python -m ttc_amd.analytics [--db PATH]
"""
import argparse
import json
import sqlite3
import sys

from . import ttc


def _median(xs):
    xs = sorted(xs)
    n = len(xs)
    return (xs[n // 2] if n % 2 else 0.5 * (xs[n // 2 - 1] + xs[n // 2])) if n else float("nan")


def summarize(db_path):
    """{'measurements': n, 'by_problem': [...], 'winners': [...], 'planner': {...}}."""
    con = sqlite3.connect(db_path)
    try:
        rows = con.execute(
            """SELECT m.measurement_id, m.dim, m.floatTypeA, m.floatTypeB, m.beta, v.bandwidth, v.config,
                      v.family, v.vecA, v.vecB, v.tileOrder
               FROM measurements m JOIN variant v ON v.measurement_id = m.measurement_id
               WHERE v.bandwidth IS NOT NULL""").fetchall()
    finally:
        con.close()
    meas = {}
    for mid, dim, ta, tb, beta, bw, cfg, fam, va, vb, tord in rows:
        m = meas.setdefault(mid, {"dim": dim, "pair": "%s->%s" % (ta, tb), "beta0": beta == 0, "variants": []})
        m["variants"].append({"bw": bw, "planner": (cfg or "").strip() in ("{}", ""), "family": fam, "vec": (va, vb),
                              "tile_order": tord})
    groups, winners = {}, {}
    gains, planner_best, with_planner = [], 0, 0
    for m in meas.values():
        best = max(m["variants"], key=lambda v: v["bw"])
        groups.setdefault((m["pair"], m["dim"]), []).append(best["bw"])
        key = (best["family"], best["vec"][0], best["vec"][1], best["tile_order"])
        winners[key] = winners.get(key, 0) + 1
        plan = [v for v in m["variants"] if v["planner"]]
        if plan:
            with_planner += 1
            p = max(plan, key=lambda v: v["bw"])
            planner_best += p["bw"] >= best["bw"]
            if p["bw"] > 0:
                gains.append(best["bw"] / p["bw"])
    by_problem = [{"pair": pair, "rank": dim, "n": len(v), "min": min(v), "median": _median(v), "max": max(v)}
                  for (pair, dim), v in sorted(groups.items())]
    win = [{"family": k[0], "vec_a": k[1], "vec_b": k[2], "tile_order": k[3], "count": n}
           for k, n in sorted(winners.items(), key=lambda kv: -kv[1])]
    return {"measurements": len(meas), "by_problem": by_problem, "winners": win,
            "planner": {"measurements": with_planner, "planner_fastest": planner_best,
                        "median_gain": _median(gains), "max_gain": max(gains) if gains else float("nan")}}


FAMILY = {1: "box", 2: "row x column", 3: "row copy"}


def render(s):
    out = ["# Tuning database summary", "", "%d measurements." % s["measurements"], "",
           "Best bandwidth found (GiB/s, the reference's unit) per type pair and fused rank:", "",
           "| types | rank | problems | min | median | max |", "|---|---:|---:|---:|---:|---:|"]
    for r in s["by_problem"]:
        out.append("| %s | %d | %d | %.1f | %.1f | %.1f |" % (r["pair"], r["rank"], r["n"], r["min"], r["median"],
                                                              r["max"]))
    out += ["", "Winning configurations:", "", "| family | vec A | vec B | tile order | problems |",
            "|---|---:|---:|---:|---:|"]
    for w in s["winners"]:
        out.append("| %s | %s | %s | %s | %d |" % (FAMILY.get(w["family"], w["family"]), w["vec_a"], w["vec_b"],
                                                   w["tile_order"], w["count"]))
    p = s["planner"]
    out += ["", "The planner's own choice was the fastest in %d of %d searches; the search gained %.3fx at the "
            "median (%.3fx at most)." % (p["planner_fastest"], p["measurements"], p["median_gain"], p["max_gain"])]
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--db", default=None, help="database path (default: TTC's location, see ttc_amd.ttc)")
    ap.add_argument("--json", action="store_true")
    a = ap.parse_args(argv)
    s = summarize(a.db or ttc.database_path())
    sys.stdout.write(json.dumps(s, indent=1) + "\n" if a.json else render(s))
    return 0


if __name__ == "__main__":
    sys.exit(main())
