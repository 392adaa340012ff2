"""Text summaries of rocprofv3 runs for the profiles/ folder, and the reader
bench.py uses for the measured HBM traffic of its roofline kernel.

Two inputs are summarised:

* a kernel trace (``rocprofv3 --kernel-trace --stats --output-format csv``):
  per-kernel launches, average duration and share of GPU time.  Arnoldi-step
  kernels enqueued past a cycle's convergence return at once; launches shorter
  than ``idle_us`` of those kernels are counted apart, as the library's own
  HIP-event timing does;
* two PMC runs (``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE``, KiB per
  dispatch) of the same command, plus the same two counters for a plain device
  copy of known size (``bench.py --calibrate-copy MIB``).  The copy's counter
  reading gives the factor from counter KiB to bytes actually moved (on this
  part FETCH_SIZE reads half the bytes a copy reads: factor 2; WRITE_SIZE
  factor 1); the kernel's readings are scaled by it.

Usage::

    python bench_profiles.py KERNEL TRACE_DIR OUT.txt "COMMAND" \\
        [--bench BENCH_LOG] [--bench-plain UNPROFILED_BENCH_LOG]
        [--pmc FETCH_DIR WRITE_DIR CAL_FETCH_DIR CAL_WRITE_DIR CAL_MIB "PMC COMMAND" [--pmc-bench PMC_BENCH_LOG]]
    python bench_profiles.py --counters OUT.txt "COMMAND" PASS_DIR [PASS_DIR ...]
    python bench_profiles.py --busy OUT.txt "COMMAND" TRACE_DIR

With --pmc-bench (the bench.py log of the PMC runs, run with --kernel KERNEL)
the counters are averaged over that run's timed-region dispatches only and
compared with their own algorithmic bytes (traffic_over_algorithmic).
--counters writes, per kernel of the step, the occupancy / VALU / LDS
counters of the given rocprofv3 --pmc runs (one counter set per run) over
the working dispatches.  --busy writes the GPU busy fraction and kernel
concurrency per 50 ms of a kernel trace (the batched ensemble).

KERNEL is the library's kernel name (``mdot``, ``maxpy_norm``, ``jacobian``,
...).  The summary ends with machine-readable ``key: value`` lines, which
:func:`read_summary` parses.
"""

import csv
import json
import os
import statistics
import sys

# library kernel name -> substring of the HIP kernel symbol
SYMBOLS = {
    "mdot": "::mdot_kernel<", "maxpy_norm": "::maxpy_norm_kernel<", "jacobian": "bracket_kernel<4, 1, double>",
    "residual": "bracket_kernel<4, 0, double>", "rfft": "::rfft_kernel<double", "irfft": "::irfft_kernel<double",
    "tridiag": "tridiag_lu_kernel<8, double>", "mdot32": "::mdot_kernel<", "maxpy_norm32": "::maxpy_norm_kernel<",
    "jacobian32": "bracket_kernel<4, 1, float>", "irfft_mdot": "::irfft_mdot_kernel<double",
    "irfft_mdot32": "::irfft_mdot_kernel<float", "jac_rfft": "::jac_rfft_kernel<4, 4, 4, double>",
    "jac_rfft32": "::jac_rfft_kernel<4, 4, 4, float>",
}
# kernels that return at once when enqueued past a cycle's convergence
EARLY = ("mdot_kernel", "maxpy_norm_kernel", "bracket_kernel<4, 1", "rfft_kernel", "tridiag_kernel",
         "tridiag_lu_kernel", "irfft_kernel", "irfft_mdot_kernel", "jac_rfft_kernel")


def _short(name):
    for p in ("vv::(anonymous namespace)::", "(anonymous namespace)::", "void "):
        name = name.replace(p, "")
    return name.split("(")[0]


def _dur_us(r):
    return (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0


def _is_double(name, kernel):
    # FP64 and FP32 instances share a symbol prefix; the template argument decides
    if kernel.endswith("32"):
        return "float" in name
    return "float" not in name


def trace_table(trace_csv, idle_us=8.0):
    rows = list(csv.DictReader(open(trace_csv)))
    by = {}
    for r in rows:
        by.setdefault(_short(r["Kernel_Name"]), []).append(_dur_us(r))
    tot = sum(sum(v) for v in by.values())
    out = ["%-58s %7s %7s %10s %10s %6s" % ("kernel", "calls", "work", "avg_us", "total_ms", "%")]
    for name, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        w = ([x for x in v if x > idle_us] or v) if any(t in name for t in EARLY) else v
        out.append("%-58s %7d %7d %10.2f %10.3f %6.1f" % (name[:58], len(v), len(w), statistics.mean(w),
                                                           sum(v) / 1e3, 100 * sum(v) / tot))
    return rows, out


def kernel_launches(rows, kernel, idle_us=8.0):
    """Durations (us) of the working launches of one library kernel, in launch order."""
    sym = SYMBOLS[kernel]
    sel = sorted((r for r in rows if sym in r["Kernel_Name"] and _is_double(r["Kernel_Name"], kernel)),
                 key=lambda r: int(r["Start_Timestamp"]))
    return [d for d in (_dur_us(r) for r in sel) if d > idle_us]


def pmc_kib(counter_dir, counter, sym, kernel=None, idle_us=8.0, big_only=False, last=None):
    """Mean counter value (KiB) over the working dispatches whose symbol contains
    sym (big_only: only dispatches reading at least half the largest, i.e. the
    calibration copies themselves; last: only the last `last` working
    dispatches, in dispatch order -- the timed region of the bench run)."""
    path = os.path.join(counter_dir, "run_counter_collection.csv")
    vals = []
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter or sym not in r["Kernel_Name"]:
            continue
        if kernel is not None and not _is_double(r["Kernel_Name"], kernel):
            continue
        if kernel is not None and _dur_us(r) <= idle_us:
            continue
        vals.append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    if not vals:
        raise ValueError("no %s dispatches of %s in %s" % (counter, sym, path))
    vals = [v for _, v in sorted(vals)]
    if big_only:
        vals = [v for v in vals if v >= 0.5 * max(vals)]
    if last:
        vals = vals[-last:]
    return statistics.mean(vals), len(vals)


def bench_roofline(log):
    """(algorithmic bytes per launch, launches) of the timed kernel in a bench.py log."""
    for ln in open(log):
        if ln.startswith("{"):
            r = json.loads(ln).get("roofline") or {}
            if "bytes_per_launch" in r:
                return r["bytes_per_launch"], int(r["launches"])
    raise ValueError("no roofline object in %s" % log)


def summarise(kernel, trace_dir, out, command, bench_log=None, pmc=None, plain_log=None):
    rows, table = trace_table(os.path.join(trace_dir, "run_kernel_trace.csv"))
    lines = ["command: " + command, "",
             "per kernel: launches; launches doing work (Arnoldi-step kernels enqueued past convergence",
             "return at once: duration <= 8 us); average duration of those; total time; share of GPU time", ""]
    lines += table
    keys = {}
    d = kernel_launches(rows, kernel)
    keys["kernel"] = kernel
    keys["rocprof_launches"] = len(d)
    keys["rocprof_avg_us"] = round(statistics.mean(d), 3)
    if bench_log:
        for ln in open(bench_log):
            if not ln.startswith("{"):
                continue
            j = json.loads(ln)
            ks = j.get("kernels", {}).get(kernel)
            lines += ["", "bench.py line of the same run (HIP events attached to the dispatches):",
                      "value: %s %s, ms_per_step %s" % (j["value"], j["unit"], j["ms_per_step"])]
            if ks:
                n = ks["launches"]
                keys["bench_launches"] = n
                keys["bench_avg_us"] = ks["avg_us"]
                keys["rocprof_avg_us_timed_region"] = round(statistics.mean(d[-n:]), 3)
                lines.append("%s: rocprof average of the last %d working launches (the timed region) %.2f us; "
                             "bench.py HIP-event average %.2f us" % (kernel, n, statistics.mean(d[-n:]),
                                                                     ks["avg_us"]))
    if plain_log:
        for ln in open(plain_log):
            if ln.startswith("{"):
                j = json.loads(ln)
                ks = j.get("kernels", {}).get(kernel)
                if ks:
                    keys["bench_avg_us_unprofiled"] = ks["avg_us"]
                    lines.append("%s: bench.py HIP-event average of the same command run without rocprofv3 "
                                 "(the per-dispatch tracing lengthens event-timed dispatches) %.2f us, value %s %s"
                                 % (kernel, ks["avg_us"], j["value"], j["unit"]))
    if pmc:
        fdir, wdir, cfdir, cwdir, cal_mib, pcmd = pmc[:6]
        sym = SYMBOLS[kernel]
        last = algo = None
        if len(pmc) > 6:
            # the PMC runs' own bench line: the algorithmic bytes of exactly the
            # dispatches whose counters are averaged (its timed region)
            algo, last = bench_roofline(pmc[6])
        f_kib, nf = pmc_kib(fdir, "FETCH_SIZE", sym, kernel, last=last)
        w_kib, nw = pmc_kib(wdir, "WRITE_SIZE", sym, kernel, last=last)
        cf_kib, _ = pmc_kib(cfdir, "FETCH_SIZE", "copyBuffer", big_only=True)
        cw_kib, _ = pmc_kib(cwdir, "WRITE_SIZE", "copyBuffer", big_only=True)
        cal_kib = float(cal_mib) * 1024.0
        ff, fw = cal_kib / cf_kib, cal_kib / cw_kib
        fetch = f_kib * ff * 1024.0
        write = w_kib * fw * 1024.0
        lines += ["", "HBM traffic (PMC, command: %s):" % pcmd,
                  "  %s FETCH_SIZE %.1f KiB/dispatch over %d dispatches, WRITE_SIZE %.1f KiB/dispatch over %d"
                  % (kernel, f_kib, nf, w_kib, nw),
                  "  copy of %s MiB reads as FETCH_SIZE %.1f KiB (factor %.3f), WRITE_SIZE %.1f KiB (factor %.3f)"
                  % (cal_mib, cf_kib, ff, cw_kib, fw),
                  "  scaled: fetch %.0f B + write %.0f B = %.0f B per dispatch" % (fetch, write, fetch + write)]
        keys.update(fetch_bytes_per_launch=round(fetch), write_bytes_per_launch=round(write),
                    traffic_bytes_per_launch=round(fetch + write), fetch_factor=round(ff, 4),
                    write_factor=round(fw, 4))
        if algo:
            lines.append("  the same %d dispatches (the PMC runs' timed region) carry %.0f algorithmic bytes each: "
                         "traffic / algorithmic = %.3f" % (last, algo, (fetch + write) / algo))
            keys.update(pmc_algorithmic_bytes_per_launch=round(algo),
                        traffic_over_algorithmic=round((fetch + write) / algo, 4))
    lines += [""] + ["%s: %s" % (k, v) for k, v in keys.items()]
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return keys


def counters_table(pass_dirs, idle_us=8.0):
    """Per kernel: the mean of every counter of the given --pmc runs over the
    working dispatches (longer than idle_us), plus the derived ratios."""
    per = {}
    for d in pass_dirs:
        for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
            if _dur_us(r) <= idle_us:
                continue
            per.setdefault(_short(r["Kernel_Name"]), {}).setdefault(r["Counter_Name"], []).append(
                float(r["Counter_Value"]))
    lines = ["%-40s %6s %6s %6s %6s %8s %8s %8s" % ("kernel (working dispatches)", "occ%", "VALU%", "LDS%",
                                                    "bank%", "wait", "VALU/wv", "LDS/wv")]
    keys = {}
    for name in sorted(per, key=lambda n: -sum(per[n].get("SQ_WAVE_CYCLES", [0]))):
        m = {c: statistics.mean(v) for c, v in per[name].items()}
        if "SQ_WAVES" not in m or m["SQ_WAVES"] < 64:
            continue
        wait = m["SQ_WAIT_INST_ANY"] / m["SQ_WAVE_CYCLES"] if m.get("SQ_WAVE_CYCLES") else float("nan")
        lines.append("%-40s %6.1f %6.1f %6.1f %6.2f %8.2f %8.0f %8.0f" % (
            name[:40], m.get("OccupancyPercent", float("nan")), m.get("VALUBusy", float("nan")),
            m.get("LdsUtil", float("nan")), m.get("LdsBankConflict", float("nan")), wait,
            m.get("SQ_INSTS_VALU", float("nan")) / m["SQ_WAVES"], m.get("SQ_INSTS_LDS", float("nan")) / m["SQ_WAVES"]))
        tag = name.split("<")[0]
        if tag not in keys:
            keys[tag] = m
    return lines, keys


def write_counters(out, command, pass_dirs):
    lines, _ = counters_table(pass_dirs)
    head = ["command: " + command, "",
            "occ%: OccupancyPercent (waves resident / wave slots, averaged over the kernel); VALU%: VALUBusy;",
            "LDS%: LdsUtil; bank%: LdsBankConflict; wait: SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES (share of a",
            "wave's life spent waiting on a dependency); VALU/wv, LDS/wv: SQ_INSTS_VALU, SQ_INSTS_LDS per wave.",
            "Averages over the working dispatches (Arnoldi-step kernels enqueued past convergence return",
            "at once and are left out: duration <= 8 us).", ""]
    with open(out, "w") as fh:
        fh.write("\n".join(head + lines) + "\n")


def busy_table(trace_csv, window_ms=50.0):
    """GPU busy fraction (union of kernel intervals) and kernel concurrency
    (summed kernel time / window) per window of a kernel trace, plus the
    share of the trace spent with one kernel at a time."""
    ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Stream_Id"])
                for r in csv.DictReader(open(trace_csv)))
    t0, t1 = ev[0][0], max(e for _, e, _ in ev)
    W = int(window_ms * 1e6)
    lines = ["%8s %8s %8s %8s %8s" % ("t_ms", "busy", "concur", "kernels", "streams")]
    for w in range(int((t1 - t0) // W) + 1):
        a, b = t0 + w * W, t0 + (w + 1) * W
        sub = sorted((max(s, a), min(e, b), st) for s, e, st in ev if e > a and s < b)
        if not sub:
            continue
        busy, cs, ce = 0, None, None
        for s, e, _ in sub:
            if ce is None or s > ce:
                if ce is not None:
                    busy += ce - cs
                cs, ce = s, e
            else:
                ce = max(ce, e)
        busy += ce - cs
        tot = sum(e - s for s, e, _ in sub)
        lines.append("%8.0f %8.2f %8.2f %8d %8d" % ((a - t0) / 1e6, busy / W, tot / W, len(sub),
                                                  len({st for _, _, st in sub})))
    return lines


def write_busy(out, command, trace_dir):
    head = ["command: " + command, "",
            "per 50 ms window of the kernel trace: busy = share of the window with at least one kernel",
            "running; concur = summed kernel time / window (kernels running at once); streams = HIP",
            "streams with a kernel in the window.", ""]
    with open(out, "w") as fh:
        fh.write("\n".join(head + busy_table(os.path.join(trace_dir, "run_kernel_trace.csv"))) + "\n")


def read_summary(path):
    """The key: value lines of a summary written by summarise()."""
    keys = {}
    if not os.path.exists(path):
        return keys
    for ln in open(path):
        if ":" in ln and not ln.startswith(" "):
            k, _, v = ln.partition(":")
            k = k.strip()
            if k and " " not in k:
                v = v.strip()
                try:
                    keys[k] = float(v) if any(c in v for c in ".e") else int(v)
                except ValueError:
                    keys[k] = v
    return keys


def main(argv):
    if argv and argv[0] == "--counters":
        write_counters(argv[1], argv[2], argv[3:])
        return
    if argv and argv[0] == "--busy":
        write_busy(argv[1], argv[2], argv[3])
        return
    kernel, trace_dir, out, command = argv[:4]
    rest = argv[4:]
    bench_log = rest[rest.index("--bench") + 1] if "--bench" in rest else None
    plain_log = rest[rest.index("--bench-plain") + 1] if "--bench-plain" in rest else None
    pmc = None
    if "--pmc" in rest:
        i = rest.index("--pmc")
        pmc = rest[i + 1:i + 7]
        if "--pmc-bench" in rest:
            pmc.append(rest[rest.index("--pmc-bench") + 1])
    keys = summarise(kernel, trace_dir, out, command, bench_log, pmc, plain_log)
    print(json.dumps(keys))


if __name__ == "__main__":
    main(sys.argv[1:])
