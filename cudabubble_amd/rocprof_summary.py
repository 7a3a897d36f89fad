"""Summaries of rocprofv3 output for the solver's kernels.

    python -m cudabubble_amd.rocprof_summary stats  KERNEL_STATS.csv  --cmd "..."
    python -m cudabubble_amd.rocprof_summary traffic --fetch F.csv --write W.csv \
        --cal-fetch CF.csv --cal-write CW.csv --n 256 --cmd "..." [--json-dir DIR --json-prec fp32]
    python -m cudabubble_amd.rocprof_summary sequence KERNEL_TRACE.csv --n 256 --prec fp32 --cmd "..."
    python -m cudabubble_amd.rocprof_summary counters A.csv B.csv --n 256 --cmd "..." --cmd "..."

`stats` turns a `--kernel-trace --stats` kernel_stats.csv into a markdown
table keyed by the solver's own kernel names (the names `vb_profile` uses).
`traffic` turns two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs into HBM bytes
per launch, each counter scaled by what it reads for a plain device copy of
known size (the `bench.py --copy-calibration` run), and optionally writes one
`traffic_<kernel>_<n>_<prec>.json` per kernel for bench.py's roofline.
`sequence` reads a `--kernel-trace` kernel_trace.csv and reports, for the
solver kernels of one size and precision, the average duration of each kernel
by the kernel that ran just before it, and the idle time between consecutive
dispatches.
`counters` joins `--pmc` runs of SQ and GRBM counters into instructions per
wave and the VALU issue fraction of each kernel of one length.
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

COPY_BYTES = 256 * 1024 * 1024     # bench.py --copy-calibration: bytes read and written per copy


def solver_name(mangled):
    """Maps a demangled rocprof kernel name to (solver kernel name, precision) or None."""
    if "k_finalize" in mangled:
        return "finalize", None
    if "k_fill_scal" in mangled:
        return "fill_scal", None
    m = re.search(r"vb::k_(\w+)<([^>]*)>", mangled)
    if not m:
        return None
    kind = m.group(1)
    args = [x.strip() for x in m.group(2).split(",")]
    if args[0] not in ("double", "float"):
        return None
    prec = "fp64" if args[0] == "double" else "fp32"
    a = args[1:] + [None] * 6
    if kind == "xa":   # <T, N, WITHF, TAB>
        return ("xa_f" if a[1] == "true" else "xa_plain"), prec
    if kind == "xrow":   # the x kernel of the ETD stages: <T, N, TAB, XP, SUMS>
        return ("xrow_sums" if a[3] == "true" else "xa"), prec
    if kind == "xsums":
        return "xa_sums", prec
    if kind == "ypass":
        return ("y_fwd" if a[1] == "-1" else "y_inv"), prec
    if kind == "zpass":
        return ("z_fwd" if a[1] == "-1" else "z_inv"), prec
    if kind == "zstage":   # <T, N, SCH, STG, TB, TAB, V, PROBE>
        if a[6] == "true":
            return "zstage4p", prec
        return "zstage%s_etdrk%s" % (a[2], a[1]) if a[1] != "4" else "zstage%s" % a[2], prec
    if kind == "zlap":
        return "z_lap", prec
    return kind, prec


def kernel_size(mangled):
    m = re.search(r"vb::k_\w+<(?:double|float)(?:, ?(?:double|float))?, ?(\d+)", mangled)
    return m.group(1) if m else "-"


def stats(path, cmd):
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["# Kernel time, rocprofv3 --kernel-trace --stats", "",
           "Command (repository root):", "", "    " + cmd, "",
           "| kernel | n | precision | calls | avg us | total ms | % | rocprof name |",
           "|---|---:|---|---:|---:|---:|---:|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        sn = solver_name(r["Name"]) or ("-", None)
        out.append("| %s | %s | %s | %s | %.1f | %.3f | %.2f | `%s` |" % (
            sn[0], kernel_size(r["Name"]), sn[1] or "-", r["Calls"], float(r["AverageNs"]) / 1e3,
            float(r["TotalDurationNs"]) / 1e6, 100 * float(r["TotalDurationNs"]) / tot, r["Name"][:90]))
    return "\n".join(out) + "\n"


def _load(path):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def _copy_kib(cal):
    vals = [v for k, vs in cal.items() if "copyBuffer" in k for v in vs]
    if not vals:
        raise SystemExit("no copy kernel in calibration file")
    return sorted(vals)[len(vals) // 2]


def traffic(fetch, write, cal_fetch, cal_write, n, cmds, json_dir=None, json_prec=None):
    f, w = _load(fetch), _load(write)
    ff = COPY_BYTES / 1024 / _copy_kib(_load(cal_fetch))
    wf = COPY_BYTES / 1024 / _copy_kib(_load(cal_write))
    rows = []
    for k in f:
        sn = solver_name(k)
        if not sn or sn[1] is None:
            continue
        fk = sum(f[k]) / len(f[k]) * ff
        wk = (sum(w[k]) / len(w[k]) * wf) if k in w else 0.0
        rows.append((sn[0], sn[1], len(f[k]), fk * 1024, wk * 1024, k))
    out = ["# HBM traffic per launch, rocprofv3 --pmc", "",
           "FETCH_SIZE scaled by %.3f and WRITE_SIZE by %.3f: the factors that turn each counter's reading "
           "for a plain %d MiB device copy into the bytes it moves." % (ff, wf, COPY_BYTES >> 20), "",
           "Commands (repository root):", ""] + ["    " + c for c in cmds] + [
           "", "| kernel | precision | launches | read MB | written MB | total MB | template arguments |",
           "|---|---|---:|---:|---:|---:|---|"]
    # the JSON of a kernel is the row of its most-launched instantiation at
    # length n (other lengths and kernel schedules appear in the table only)
    main_row = {}
    for r in rows:
        if kernel_size(r[5]) == str(n) and ((r[0], r[1]) not in main_row or r[2] > main_row[(r[0], r[1])][2]):
            main_row[(r[0], r[1])] = r
    for name, prec, cnt, rb, wb, k in sorted(rows, key=lambda r: -(r[3] + r[4])):
        targs = re.search(r"vb::k_\w+(<[^>]*>)", k)
        out.append("| %s | %s | %d | %.1f | %.1f | %.1f | `%s` |" % (name, prec, cnt, rb / 1e6, wb / 1e6, (rb + wb) / 1e6,
                                                                  targs.group(1) if targs else "-"))
        if json_dir and (json_prec is None or prec == json_prec) and main_row.get((name, prec), (0,) * 6)[5] == k:
            with open(os.path.join(json_dir, "traffic_%s_%d_%s.json" % (name, n, prec)), "w") as fh:
                json.dump({"kernel": name, "n": n, "precision": prec, "launches": cnt,
                           "read_bytes_per_launch": rb, "write_bytes_per_launch": wb,
                           "bytes_per_launch": rb + wb, "fetch_factor": ff, "write_factor": wf,
                           "rocprof_name": k}, fh, indent=1)
    return "\n".join(out) + "\n"


def sequence(path, n, prec, cmd):
    ev = []
    for r in csv.DictReader(open(path)):
        k = r["Kernel_Name"]
        nm = solver_name(k)
        ok = nm is not None and kernel_size(k) == str(n) and nm[1] == prec
        ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), nm[0] if ok else None))
    ev.sort()
    by = collections.defaultdict(list)
    gaps = []
    for (s0, e0, k0), (s1, e1, k1) in zip(ev, ev[1:]):
        if k0 is None or k1 is None:
            continue
        by[(k1, k0)].append((e1 - s1) * 1e-3)
        gaps.append((s1 - e0) * 1e-3)
    out = ["# Kernel durations by predecessor, rocprofv3 --kernel-trace (%d^3, %s)" % (n, prec), "",
           "Command (repository root):", "", "    " + cmd, "",
           "| kernel | after | launches | avg us |", "|---|---|---:|---:|"]
    for (k1, k0), v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        out.append("| %s | %s | %d | %.1f |" % (k1, k0, len(v), sum(v) / len(v)))
    if gaps:
        gs = sorted(gaps)
        out += ["", "Idle time between consecutive dispatches of these kernels: %d gaps, mean %.2f us, "
                "median %.2f us, 99th percentile %.2f us." % (len(gs), sum(gs) / len(gs), gs[len(gs) // 2],
                                                               gs[min(len(gs) - 1, int(0.99 * len(gs)))])]
    return "\n".join(out) + "\n"


N_SIMD = 1024      # MI355X: 256 CUs x 4 SIMDs
N_XCD = 8          # GRBM_GUI_ACTIVE is summed over the 8 XCDs


def _pmc_means(paths):
    """{(kernel name, counter): mean per dispatch} over the dispatches of each kernel."""
    acc = collections.defaultdict(list)
    for p in paths:
        for r in csv.DictReader(open(p)):
            acc[(r["Kernel_Name"], r["Counter_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}


def counters(paths, n, cmds):
    """Per kernel of length n, both precisions: instructions per wave by kind,
    the VALU issue fraction and the LDS bank-conflict share.  A wave64 VALU
    instruction (FP32 unpacked or FP64) holds its 16-lane SIMD for 4 cycles, so
    4 SQ_INSTS_VALU / (SIMDs x kernel cycles) is the fraction of the SIMDs'
    cycles spent issuing VALU work; kernel cycles are GRBM_GUI_ACTIVE / XCDs."""
    m, cnt = _pmc_means(paths)
    kernels = sorted({k for k, _ in m})
    rows = []
    for k in kernels:
        sn = solver_name(k)
        if not sn or sn[1] is None or kernel_size(k) != str(n):
            continue
        g = lambda c: m.get((k, c))
        waves, valu, grbm = g("SQ_WAVES"), g("SQ_INSTS_VALU"), g("GRBM_GUI_ACTIVE")
        if not waves or valu is None:
            continue
        cyc = grbm / N_XCD if grbm else None
        per = lambda c: (g(c) / waves) if g(c) is not None else None
        rows.append({
            "kernel": sn[0], "prec": sn[1], "launches": cnt.get((k, "SQ_WAVES"), 0), "waves": waves,
            "cycles": cyc, "valu_pw": valu / waves, "lds_pw": per("SQ_INSTS_LDS"),
            "vmrd_pw": per("SQ_INSTS_VMEM_RD"), "vmwr_pw": per("SQ_INSTS_VMEM_WR"),
            "valu_frac": 4 * valu / (N_SIMD * cyc) if cyc else None,
            "conf": g("SQ_LDS_BANK_CONFLICT"),
            "wait": (g("SQ_WAIT_INST_ANY") / g("SQ_WAVE_CYCLES")) if g("SQ_WAIT_INST_ANY") and g("SQ_WAVE_CYCLES") else None,
            "name": k})
    f = lambda x, s: ("-" if x is None else s % x)
    out = ["# Instruction mix and VALU issue per launch, rocprofv3 --pmc (%d^3)" % n, "",
           "Commands (repository root):", ""] + ["    " + c for c in cmds] + [
           "", "Means over each kernel's launches.  VALU issue = 4 SQ_INSTS_VALU / (%d SIMDs x "
           "GRBM_GUI_ACTIVE / %d XCDs): the fraction of SIMD cycles that issued VALU work (a wave64 "
           "VALU instruction, FP32 unpacked or FP64, holds a 16-lane SIMD for 4 cycles).  Wait = "
           "SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES.  LDS conflict = SQ_LDS_BANK_CONFLICT (cycles, summed "
           "over CUs) per launch." % (N_SIMD, N_XCD), "",
           "| kernel | precision | waves | kernel cycles | VALU / wave | LDS / wave | VMEM rd / wave | "
           "VMEM wr / wave | VALU issue | wait | LDS conflict |",
           "|---|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: (r["kernel"], r["prec"])):
        out.append("| %s | %s | %d | %s | %.0f | %s | %s | %s | %s | %s | %s |" % (
            r["kernel"], r["prec"], r["waves"], f(r["cycles"], "%.0f"), r["valu_pw"], f(r["lds_pw"], "%.0f"),
            f(r["vmrd_pw"], "%.0f"), f(r["vmwr_pw"], "%.0f"), f(r["valu_frac"], "%.2f"), f(r["wait"], "%.2f"),
            f(r["conf"], "%.3g")))
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--cal-fetch", "--cal-write"):
        t.add_argument(a, required=True)
    t.add_argument("--n", type=int, required=True)
    t.add_argument("--cmd", action="append", required=True)
    t.add_argument("--json-dir")
    t.add_argument("--json-prec", help="write JSON only for this precision (fp32 / fp64)")
    q = sub.add_parser("sequence")
    q.add_argument("csv")
    q.add_argument("--n", type=int, required=True)
    q.add_argument("--prec", default="fp32")
    q.add_argument("--cmd", required=True)
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--n", type=int, required=True)
    c.add_argument("--cmd", action="append", required=True)
    a = ap.parse_args(argv)
    if a.what == "counters":
        sys.stdout.write(counters(a.csv, a.n, a.cmd))
    elif a.what == "sequence":
        sys.stdout.write(sequence(a.csv, a.n, a.prec, a.cmd))
    elif a.what == "stats":
        sys.stdout.write(stats(a.csv, a.cmd))
    else:
        sys.stdout.write(traffic(a.fetch, a.write, a.cal_fetch, a.cal_write, a.n, a.cmd, a.json_dir, a.json_prec))


if __name__ == "__main__":
    main()
