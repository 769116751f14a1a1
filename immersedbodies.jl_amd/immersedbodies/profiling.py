"""Per-step kernel breakdown and counter summaries from rocprofv3 databases.

    python -m immersedbodies.profiling RESULTS_DB [--steps N] [--marker NAME]
    python -m immersedbodies.profiling workload --kernel K --precision P   (program to profile)
    python -m immersedbodies.profiling pmc RESULTS_DB --match NAME         (--pmc summary)
    python -m immersedbodies.profiling trace KERNEL_TRACE_CSV --match NAME (timing-loop mean)

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
    # a step's marker is followed by the step's first Helmholtz-solve kernel, a sine
    # contraction pass or a Gauss-Seidel launch (isolated marker launches, such as a
    # kernel-timing loop, are not step boundaries)
    marks = [i for i, r in enumerate(rows)
             if marker in r[0] and i + 1 < len(rows)
             and ("dst_pass" in rows[i + 1][0] or "gs_helmholtz" in rows[i + 1][0])]
    if len(marks) < 2:
        raise ValueError("fewer than two marker dispatches")
    sel = marks[-(steps + 1):]
    n = len(sel) - 1
    per = defaultdict(lambda: [0, 0.0])
    busy = 0.0
    for i in range(sel[0], sel[-1]):
        name, s, e = rows[i]
        d = (e - s) * 1e-3  # ns -> us
        short = name.replace("(anonymous namespace)::", "").split("(")[0]
        short = short.replace("void ", "").replace("ibpm::", "")
        per[short][0] += 1
        per[short][1] += d
        busy += d
    wall = (rows[sel[-1]][1] - rows[sel[0]][1]) * 1e-3
    table = sorted(((k, c / n, t / n) for k, (c, t) in per.items()), key=lambda x: -x[2])
    return n, wall / n, busy / n, table


def workload(kernel: int, precision: str, reps: int, copy_mib: int = 2048, levels: int = 5,
             tiles: int = -1):
    """Launch one kernel of the 1024^2 cylinder step ``reps`` times (ibpm_time_kernel), then
    one plain device-to-device copy of ``copy_mib`` MiB whose byte count is known: the
    reference for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    import os
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(os.path.dirname(here))
    sys.path.insert(0, os.path.join(root, "tests"))
    import cases  # workload builders (synthetic inputs)
    import immersedbodies as ib
    from immersedbodies import _lib

    s = ib.create_solver(cases.cylinder_1024(nlevel=levels), precision=precision, dst_tiles=tiles)
    s.advance(2)
    ms, fl, by = C.c_double(), C.c_double(), C.c_double()
    _lib.check(_lib.load().ibpm_time_kernel(s._h, kernel, reps, C.byref(ms), C.byref(fl),
                                            C.byref(by)))
    print(f"kernel {kernel} {precision}: avg {ms.value:.5f} ms, flops {fl.value:.6g}, "
          f"min bytes {by.value:.6g}", flush=True)
    s.close()
    # reference copy: one device-to-device hipMemcpy (the runtime's copy kernel) of a source
    # written before a 1 GiB flush buffer, so that neither L2 nor the 256 MB last-level
    # cache still holds it
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(copy_mib * (1 << 20))
    fbytes = C.c_size_t(1 << 30)
    a, b, f = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for ptr, nb in ((a, nbytes), (b, nbytes), (f, fbytes)):
        if hip.hipMalloc(C.byref(ptr), nb) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, nbytes)
    hip.hipMemset(f, 2, fbytes)
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    for ptr in (a, b, f):
        hip.hipFree(ptr)
    print(f"copy: {copy_mib} MiB read, {copy_mib} MiB written", flush=True)


def pmc_summary(db_path: str, match: str, last: int = 11, copy_match: str = "copyBuffer"):
    """From a rocprofv3 --pmc database: the mean counter value (KiB) over the last ``last``
    dispatches of the kernel whose name contains ``match`` (those with the grid of the
    final dispatch: the timing loop of ``workload``), and the largest value on a dispatch
    of the reference copy kernel."""
    db = sqlite3.connect(db_path)
    rows = list(db.execute("select dispatch_id, kernel_name, grid_size, counter_name, value "
                           "from counters_collection order by dispatch_id"))
    per = defaultdict(lambda: defaultdict(float))  # counter -> dispatch -> value
    meta = {}
    for d, name, grid, counter, value in rows:
        per[counter][d] += value
        meta[d] = (name, grid)
    out = {}
    for counter, disp in per.items():
        ks = [d for d in sorted(disp) if match in meta[d][0]]
        if ks:
            g = meta[ks[-1]][1]
            ks = [d for d in ks if meta[d][1] == g][-last:]
        cv = [disp[d] for d in disp if copy_match in meta[d][0]]
        out[counter] = (sum(disp[d] for d in ks) / len(ks) if ks else None, len(ks),
                        meta[ks[-1]][0] if ks else None, max(cv) if cv else None)
    return out


def trace_tail(csv_path: str, match: str, last: int = 200):
    """From a rocprofv3 --kernel-trace CSV: mean and spread of the durations (us) of the last
    ``last`` dispatches of the kernel whose name contains ``match`` that share the grid of
    its final dispatch (the timing loop that follows the steps in bench.py)."""
    import csv

    rows = []
    with open(csv_path) as fh:
        for r in csv.DictReader(fh):
            if match in r["Kernel_Name"]:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size_X"]),
                             (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3))
    rows.sort()
    g, name = rows[-1][2], rows[-1][1]
    d = [r[3] for r in rows if r[2] == g and r[1] == name][-last:]
    mean = sum(d) / len(d)
    return name, g, len(d), mean, min(d), max(d)


def main():
    import sys

    if len(sys.argv) > 1 and sys.argv[1] == "trace":
        ap = argparse.ArgumentParser()
        ap.add_argument("cmd")
        ap.add_argument("csv")
        ap.add_argument("--match", default="dst_pass")
        ap.add_argument("--last", type=int, default=200)
        a = ap.parse_args()
        name, g, n, mean, lo, hi = trace_tail(a.csv, a.match, a.last)
        print(f"`{name}` grid {g}: last {n} dispatches mean {mean:.2f} us "
              f"(min {lo:.2f}, max {hi:.2f})")
        return

    if len(sys.argv) > 1 and sys.argv[1] == "workload":
        ap = argparse.ArgumentParser()
        ap.add_argument("cmd")
        ap.add_argument("--kernel", type=int, default=0)
        ap.add_argument("--precision", default="fp64")
        ap.add_argument("--reps", type=int, default=20)
        ap.add_argument("--copy-mib", type=int, default=2048)
        ap.add_argument("--tiles", type=int, default=-1, help="fixed tile variant, -1 autotuned")
        a = ap.parse_args()
        workload(a.kernel, a.precision, a.reps, a.copy_mib, tiles=a.tiles)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "pmc":
        ap = argparse.ArgumentParser()
        ap.add_argument("cmd")
        ap.add_argument("db")
        ap.add_argument("--match", default="dst_pass")
        ap.add_argument("--last", type=int, default=11)
        ap.add_argument("--copy-mib", type=int, default=2048)
        a = ap.parse_args()
        res = pmc_summary(a.db, a.match, a.last)
        for counter, (mean, n, kname, copyv) in res.items():
            scale = a.copy_mib * 1024.0 / copyv if copyv else None
            print(f"{counter}: `{kname}`: {mean:.1f} KiB per dispatch (mean of {n}); copy of "
                  f"{a.copy_mib} MiB: {copyv:.1f} KiB, scale {scale:.4f}; scaled "
                  f"{mean * scale / 1024:.2f} MiB per dispatch")
        return
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
