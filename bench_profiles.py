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
        [--pmc FETCH_DIR WRITE_DIR CAL_FETCH_DIR CAL_WRITE_DIR CAL_MIB "PMC COMMAND"]

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
    "mdot": "mdot_kernel<", "maxpy_norm": "maxpy_norm_kernel<", "jacobian": "bracket_kernel<4, 1, double>",
    "residual": "bracket_kernel<4, 0, double>", "rfft": "rfft_kernel<double", "irfft": "irfft_kernel<double",
    "tridiag": "tridiag_kernel<8, double, false>", "mdot32": "mdot_kernel<", "maxpy_norm32": "maxpy_norm_kernel<",
    "jacobian32": "bracket_kernel<4, 1, float>",
}
# kernels that return at once when enqueued past a cycle's convergence
EARLY = ("mdot_kernel", "maxpy_norm_kernel", "bracket_kernel<4, 1", "rfft_kernel", "tridiag_kernel",
         "irfft_kernel", "gmres_update_kernel", "mdot_finish_kernel")


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


def pmc_kib(counter_dir, counter, sym, kernel=None, idle_us=8.0, big_only=False):
    """Mean counter value (KiB) over the working dispatches whose symbol contains
    sym (big_only: only dispatches reading at least half the largest, i.e. the
    calibration copies themselves)."""
    path = os.path.join(counter_dir, "run_counter_collection.csv")
    vals = []
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter or sym not in r["Kernel_Name"]:
            continue
        if kernel is not None and not _is_double(r["Kernel_Name"], kernel):
            continue
        if kernel is not None and _dur_us(r) <= idle_us:
            continue
        vals.append(float(r["Counter_Value"]))
    if not vals:
        raise ValueError("no %s dispatches of %s in %s" % (counter, sym, path))
    if big_only:
        vals = [v for v in vals if v >= 0.5 * max(vals)]
    return statistics.mean(vals), len(vals)


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
        fdir, wdir, cfdir, cwdir, cal_mib, pcmd = pmc
        sym = SYMBOLS[kernel]
        f_kib, nf = pmc_kib(fdir, "FETCH_SIZE", sym, kernel)
        w_kib, nw = pmc_kib(wdir, "WRITE_SIZE", sym, kernel)
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
    lines += [""] + ["%s: %s" % (k, v) for k, v in keys.items()]
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return keys


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
    kernel, trace_dir, out, command = argv[:4]
    rest = argv[4:]
    bench_log = rest[rest.index("--bench") + 1] if "--bench" in rest else None
    plain_log = rest[rest.index("--bench-plain") + 1] if "--bench-plain" in rest else None
    pmc = None
    if "--pmc" in rest:
        i = rest.index("--pmc")
        pmc = rest[i + 1:i + 7]
    keys = summarise(kernel, trace_dir, out, command, bench_log, pmc, plain_log)
    print(json.dumps(keys))


if __name__ == "__main__":
    main(sys.argv[1:])
