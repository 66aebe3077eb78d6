"""Summaries of rocprofv3 output for profiles/: per-kernel durations from a --kernel-trace
--stats run, and HBM traffic per dispatch from --pmc FETCH_SIZE and WRITE_SIZE runs scaled by
what the same counters read for a plain device copy of known size.

    python -m eden_amd.profiling stats <run_kernel_stats.csv> <title> <command...>
    python -m eden_amd.profiling traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv>
                                         <copy MiB> <kernel-substring> <workload> <command...>
    python -m eden_amd.profiling counters <counters.csv> <title> <workload> <CUs per XCD> <command...>

The traffic summary ends with one line `traffic_bytes_per_launch: <kernel> | <workload> |
<bytes>` per kernel whose name contains the substring; bench.py reads the one of its dominant
kernel and workload for its roofline's `traffic`. The counters summary gives the mean per
dispatch of every counter in a --pmc run, per kernel, with the shares that the SQ counters
imply (LDS bank-conflict cycles over LDS-active cycles, waiting over wave cycles, and with
GRBM_GUI_ACTIVE the LDS utilisation: LDS-active cycles over the CU cycles of the dispatch,
the utilisation formula of AMD's rocprof-compute; GRBM_GUI_ACTIVE is summed over the XCDs,
so the CU cycles are GRBM_GUI_ACTIVE x CUs per XCD), and ends with one line
`counters_per_launch: <kernel> | <workload> | NAME=value;...` per kernel (with LDS_UTIL, the
utilisation), which bench.py reads for its roofline's `lds` entry.
"""
import csv
import re
import sys


def _short(name):
    name = re.sub(r"\(.*$", "", name)
    return name.replace("void ", "").replace("edenhip::", "").strip()


def kernel_stats(path, title, command):
    rows = list(csv.DictReader(open(path)))
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             "| kernel | calls | total ms | average ms | % |", "|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |")
    # per-dispatch durations of this library's kernels, from the kernel trace beside the stats
    trace = path.replace("kernel_stats.csv", "kernel_trace.csv")
    try:
        disp = [r for r in csv.DictReader(open(trace)) if "edenhip" in r["Kernel_Name"]]
    except OSError:
        disp = []
    if disp:
        lines += ["", "Dispatches in order of start (ms):", "", "| kernel | start | duration |", "|---|---|---|"]
        t0 = min(int(r["Start_Timestamp"]) for r in disp)
        for r in sorted(disp, key=lambda r: int(r["Start_Timestamp"])):
            st, en = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            lines.append(f"| {_short(r['Kernel_Name'])} | {(st - t0) / 1e6:.1f} | {(en - st) / 1e6:.3f} |")
    return "\n".join(lines) + "\n"


def _per_dispatch(path, counter):
    out = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = (r["Kernel_Name"], r["Dispatch_Id"])
        out[k] = out.get(k, 0.0) + float(r["Counter_Value"])
    return out


def traffic(fetch, write, cfetch, cwrite, copy_mib, kernel_sub, workload, command):
    nbytes = copy_mib * 1024.0   # KiB the copy reads, and writes
    f_copy = [v for (n, _), v in _per_dispatch(cfetch, "FETCH_SIZE").items() if "copyBuffer" in n]
    w_copy = [v for (n, _), v in _per_dispatch(cwrite, "WRITE_SIZE").items() if "copyBuffer" in n]
    f_copy = max(f_copy)
    w_copy = max(w_copy)
    kf, kw = nbytes / f_copy, nbytes / w_copy
    def mean_per_kernel(d):
        acc = {}
        for (n, _), v in d.items():
            if "edenhip" in n:
                t, c = acc.get(n, (0.0, 0))
                acc[n] = (t + v, c + 1)
        return {n: t / c for n, (t, c) in acc.items()}, {n: c for n, (t, c) in acc.items()}

    fk, nf = mean_per_kernel(_per_dispatch(fetch, "FETCH_SIZE"))
    wk, _ = mean_per_kernel(_per_dispatch(write, "WRITE_SIZE"))
    lines = ["# HBM traffic per dispatch, " + workload, "", "Commands (FETCH_SIZE run, then the same with WRITE_SIZE; the "
             "calibration copy with each counter):", "", "```", command, "```", "",
             f"Calibration: a plain {copy_mib} MiB device copy reads FETCH_SIZE {f_copy:.0f} KiB and WRITE_SIZE "
             f"{w_copy:.0f} KiB for {nbytes:.0f} KiB read and written: factors {kf:.3f} (fetch), {kw:.3f} (write).", "",
             "Per kernel: the mean over its dispatches.", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic bytes (calibrated) |",
             "|---|---|---|---|---|"]
    chosen = []
    for n in sorted(fk):
        t = (fk[n] * kf + wk.get(n, 0.0) * kw) * 1024.0
        lines.append(f"| {_short(n)} | {nf[n]} | {fk[n]:.0f} | {wk.get(n, 0.0):.0f} | {t:.4g} |")
        if kernel_sub in n:
            chosen.append((_short(n), t))
    lines.append("")
    for k, t in chosen:
        lines.append(f"traffic_bytes_per_launch: {k} | {workload} | {t:.6g}")
    return "\n".join(lines) + "\n"


def counters(path, title, workload, cus_per_xcd, command):
    acc, disp = {}, {}
    for r in csv.DictReader(open(path)):
        n = r["Kernel_Name"]
        if "edenhip" not in n:
            continue
        d = acc.setdefault(n, {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        disp.setdefault(n, set()).add(r["Dispatch_Id"])
    names = sorted({c for d in acc.values() for c in d})
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             "Mean per dispatch (summed over the device's shader engines).", "",
             "| kernel | dispatches | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
    for n in sorted(acc):
        k = len(disp[n])
        lines.append(f"| {_short(n)} | {k} | " + " | ".join(f"{acc[n].get(c, 0.0) / k:.4g}" for c in names) + " |")
    share = [("SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE", "LDS bank-conflict cycles / LDS-active cycles"),
             ("SQ_WAIT_ANY", "SQ_WAVE_CYCLES", "wave cycles waiting / wave cycles"),
             ("SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES", "wave cycles waiting on LDS instructions / wave cycles"),
             ("SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "VALU-active cycles / wave cycles")]
    rows = [(a, b, t) for a, b, t in share if a in names and b in names]
    util = "SQ_LDS_IDX_ACTIVE" in names and "GRBM_GUI_ACTIVE" in names
    if rows or util:
        head = [t for _, _, t in rows] + ([f"LDS utilisation (LDS-active cycles / (GRBM_GUI_ACTIVE x {cus_per_xcd} "
                                          f"CUs per XCD))"] if util else [])
        lines += ["", "| kernel | " + " | ".join(head) + " |", "|---|" + "---|" * len(head)]
        for n in sorted(acc):
            d = acc[n]
            cells = [f"{d[a] / d[b]:.3f}" if d[b] else "-" for a, b, _ in rows]
            if util:
                d["LDS_UTIL"] = d["SQ_LDS_IDX_ACTIVE"] / (d["GRBM_GUI_ACTIVE"] * cus_per_xcd) if d["GRBM_GUI_ACTIVE"] else 0.0
                cells.append(f"{d['LDS_UTIL']:.3f}")
            lines.append(f"| {_short(n)} | " + " | ".join(cells) + " |")
    lines.append("")
    for n in sorted(acc):
        k = len(disp[n])
        vals = [f"{c}={acc[n].get(c, 0.0) / k:.6g}" for c in names]
        if "LDS_UTIL" in acc[n]:
            vals.append(f"LDS_UTIL={acc[n]['LDS_UTIL']:.6g}")
        lines.append(f"counters_per_launch: {_short(n)} | {workload} | " + ";".join(vals))
    return "\n".join(lines) + "\n"


def read_counters(path, kernel, workload):
    """Mean counter values per launch recorded in a counters summary, as a dict, or None."""
    try:
        for line in open(path):
            if line.startswith("counters_per_launch:"):
                k, w, v = [x.strip() for x in line.split(":", 1)[1].split("|")]
                if k == kernel and w == workload:
                    return {a: float(b) for a, b in (x.split("=") for x in v.split(";") if x)}
    except OSError:
        return None
    return None


def read_traffic(path, kernel, workload):
    """Calibrated bytes per launch recorded in a traffic summary, or None."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                k, w, v = [x.strip() for x in line.split(":", 1)[1].split("|")]
                if k == kernel and w == workload:
                    return float(v)
    except OSError:
        return None
    return None


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        sys.stdout.write(kernel_stats(sys.argv[2], sys.argv[3], " ".join(sys.argv[4:])))
    elif sys.argv[1] == "traffic":
        a = sys.argv[2:]
        sys.stdout.write(traffic(a[0], a[1], a[2], a[3], int(a[4]), a[5], a[6], " ".join(a[7:])))
    elif sys.argv[1] == "counters":
        sys.stdout.write(counters(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]), " ".join(sys.argv[6:])))
