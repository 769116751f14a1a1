"""Per-step kernel breakdown from a rocprofv3 database (``--kernel-trace`` output).

    python -m immersedbodies.profiling RESULTS_DB [--steps N] [--marker NAME]

Kernel dispatches are ordered by start time and cut into steps at each dispatch of the
marker kernel (the fused nonlinear/right-hand-side kernel opens every time step). The
last N complete steps are aggregated: calls and GPU time per kernel per step, the GPU
busy time per step, and the wall time per step between markers.
"""
from __future__ import annotations

import argparse
import sqlite3
from collections import defaultdict


def step_breakdown(db_path: str, steps: int = 20, marker: str = "nonlin_rhs_kernel"):
    db = sqlite3.connect(db_path)
    rows = list(db.execute("select name, start, end from kernels order by start"))
    # a step's marker is followed by the step's first sine-contraction pass (isolated
    # marker launches, such as a kernel-timing loop, are not step boundaries)
    marks = [i for i, r in enumerate(rows)
             if marker in r[0] and i + 1 < len(rows) and "dst_pass" in rows[i + 1][0]]
    if len(marks) < 2:
        raise ValueError("fewer than two marker dispatches")
    sel = marks[-(steps + 1):]
    n = len(sel) - 1
    per = defaultdict(lambda: [0, 0.0])
    busy = 0.0
    for i in range(sel[0], sel[-1]):
        name, s, e = rows[i]
        d = (e - s) * 1e-3  # ns -> us
        short = name.split("(")[0].replace("void ", "").replace("ibpm::", "")
        per[short][0] += 1
        per[short][1] += d
        busy += d
    wall = (rows[sel[-1]][1] - rows[sel[0]][1]) * 1e-3
    table = sorted(((k, c / n, t / n) for k, (c, t) in per.items()), key=lambda x: -x[2])
    return n, wall / n, busy / n, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--marker", default="nonlin_rhs_kernel")
    a = ap.parse_args()
    n, wall, busy, table = step_breakdown(a.db, a.steps, a.marker)
    print(f"steps aggregated: {n}; wall per step {wall:.1f} us; kernel time per step {busy:.1f} us\n")
    print("| kernel | calls/step | us/step | % of kernel time |\n|---|---|---|---|")
    for k, c, t in table:
        print(f"| `{k}` | {c:.1f} | {t:.1f} | {100 * t / busy:.1f} |")


if __name__ == "__main__":
    main()
