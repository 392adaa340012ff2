#!/usr/bin/env python3
"""Throughput of the implicit Vlasov-Poisson timestep on one MI355X.

Workload (default): two-stream instability, nx x nv = 1024 x 2048,
x in [0, 2 pi/0.2), v in [-10, 10], f0 = (1 + 0.001 cos 0.2x) 1/2 [M(v-2.4) +
M(v+2.4)], Arakawa J4, dt = 0.05, FP64, kinetic tensor preconditioner,
GMRES(30) rtol 1e-10, Picard atol 1e-10 (at most 20 iterations).  A "step"
is one full implicit timestep: Picard loop of SNES(ksponly) + GMRES on the
preconditioned Jacobian, moments and Poisson after every solve.

Timing: W untimed warm-up steps, then 3 windows (--repeats) of K steps, each
between two torch.cuda.synchronize() calls.  value = K / (median window
time)  (timesteps/s) of the plain FP64 form.  The mixed-precision form (FP32 GMRES cycles + FP32
preconditioner for the Picard solves that start from a residual below
snes_atol * 2^20, inside the FP64 Picard residual / update / Poisson loop) is then timed the same way from
the same initial state and reported under "mixed", next to FP64 at the same
linear tolerance and with its distance to the FP64 run after --compare-steps.
The batched form (256 independent instances with uneven grids, bucketed by
shape into batched contexts, one captured graph replay per Newton step,
autotuned launch shapes) is reported under "batched": aggregate
timesteps/s over all instances next to the same instances run one after
another through the single-instance path.
roofline: the dominant kernel's algorithmic bytes / its HIP-event time on
the library stream over the timed region.  cpu_baseline: the C (OpenMP)
restatement of the same timestep from oracle/, on a bounded sample.

Prints ONE JSON line.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = ("jacobian", "residual", "rfft", "tridiag", "irfft", "mdot", "maxpy_norm",
           "jacobian32", "rfft32", "tridiag32", "irfft32", "mdot32", "maxpy_norm32", "irfft_mdot", "irfft_mdot32",
           "jac_rfft", "jac_rfft32")


def workloads():
    import numpy as np
    from vivlasov1d_amd.core import landau, twostream
    return {
        "twostream_1024x2048_J4": dict(nx=1024, nv=2048, L=2 * np.pi / 0.2, vmax=10.0, ht=0.05,
                                       bracket="arakawaj4", profile=twostream(0.2, 0.001),
                                       desc="two-stream k=0.2 amp=1e-3 vd=2.4"),
        "landau_256x512_J4": dict(nx=256, nv=512, L=4 * np.pi, vmax=8.0, ht=0.05, bracket="arakawaj4",
                                  profile=landau(0.5, 0.5), desc="nonlinear Landau k=0.5 amp=0.5"),
        "landau_64x128_J4": dict(nx=64, nv=128, L=4 * np.pi, vmax=10.0, ht=0.1, bracket="arakawaj4",
                                 profile=landau(0.5, 0.01), desc="linear Landau k=0.5 amp=0.01"),
    }


def make_cfg(w, nt):
    from vivlasov1d_amd.core import Config
    return Config({
        "grid": {"nx": w["nx"], "nv": w["nv"], "L": w["L"], "vmax": w["vmax"], "ht": w["ht"], "nt": nt},
        "solver": {"method": "matrixfree", "type": "nonlinear", "mode": "split", "timestepping": "mp",
                   "poisson_bracket": w["bracket"], "laplace_operator": "CFD4",
                   "preconditioner_type": "tensor", "preconditioner_scheme": "tensor_kinetic",
                   "petsc_ksp_rtol": 1e-10, "petsc_ksp_atol": 1e-50, "petsc_ksp_max_iter": 1000,
                   "petsc_snes_atol": 1e-10, "petsc_snes_max_iter": 20},
        "initial_data": {"charge": -1.0, "mass": 1.0},
    })


def step_bytes(N, rec, restart=30):
    """Algorithmic HBM bytes of one timestep from its Picard / GMRES counts."""
    total = 24.0
    for p in rec["picard"]:
        m = p["gmres"]
        s = 16.0
        for k in range(1, m + 1):
            kk = ((k - 1) % restart) + 1
            s += 32.0 + 8.0 * (2 * kk + 3)
        s += 8.0 * (m + 2) + 8.0 + 24.0
        total += s
    return N * total


def cpu_baseline(w, sample_steps, threads):
    """C restatement of the timestep (oracle/c), timed on the host cores."""
    from oracle import cbaseline
    return cbaseline.time_sample(w, sample_steps, threads)


def timed_run(w, args, precision, kernel, refine="newton", ksp_rtol=None, repeats=None, fp64_above="auto",
              every=1):
    """W warm-up steps, then `repeats` windows of K timed steps each; returns
    the run, the median window time, the records of all timed steps and the
    per-kernel event statistics (summed over the windows)."""
    import torch
    from vivlasov1d_amd.runner import Runscript
    repeats = args.repeats if repeats is None else repeats
    nt = args.warmup + repeats * args.steps
    run = Runscript(make_cfg(w, nt), profile=w["profile"], precision=precision, inner_rtol=args.inner_rtol,
                    refine=refine, fp64_above=fp64_above)
    if ksp_rtol is not None:
        run.snes.getKSP().setTolerances(rtol=ksp_rtol)
    run.inv0 = run.invariants()
    ctx = run.ctx
    if args.no_fusion:
        ctx.set_fusion(False)
    for _ in range(args.warmup):
        run.step()
    torch.cuda.synchronize()
    if kernel == "all":
        ctx.profile(None, every=every)
    elif kernel not in (None, "none"):
        ctx.profile(tuple(kernel.split(",")), every=every)
    nrec0 = len(run.history)
    windows = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run.step()
        torch.cuda.synchronize()
        windows.append(time.perf_counter() - t0)
    ctx.profile(())
    kstats = {}
    for k in KERNELS:
        ms, n, by = ctx.profile_read(k)
        if n:
            kstats[k] = dict(ms=ms, launches=n, bytes=by, avg_us=1e3 * ms / n, gbs=by / (ms * 1e-3) / 1e9)
    run.windows = windows
    return run, sorted(windows)[len(windows) // 2], run.history[nrec0:], kstats


def bracket_config2(args):
    """The metric's bracket HBM GB/s at the nonlinear-Landau workload
    (256x512, v in [-8, 8], amplitude 0.5, dt = 0.05) for J1, J2 and J4: the
    residual (24 B/point), the Arnoldi step's J u fused with the rFFT
    (8 B/point in + the spectrum out) and the stand-alone Jacobian-apply of the
    restart residuals (16 B/point), HIP-event timed over a few timesteps."""
    out = {}
    base = workloads()["landau_256x512_J4"]
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.repeats = 5, 2, 1
    for kind in ("arakawaj1", "arakawaj2", "arakawaj4"):
        w = dict(base, bracket=kind)
        brun, _, _, bk = timed_run(w, sub, "double", "jac_rfft,residual,jacobian")
        del brun
        out[kind] = {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"], "GB/s": round(v["gbs"], 1),
                         "frac_hbm_peak": round(v["gbs"] / HBM_PEAK_GBS, 4)} for k, v in bk.items()}
    return out


def _its(recs):
    import numpy as np
    gm = [p["gmres"] for r in recs for p in r["picard"]]
    return {"picard_per_step": float(np.mean([len(r["picard"]) for r in recs])),
            "gmres_per_solve": float(np.mean(gm)), "gmres_per_step": float(sum(gm)) / max(1, len(recs))}


def mixed_figures(w, args, run64, value64):
    """The mixed-precision form (FP32 GMRES cycles and P^-1 inside the FP64
    Picard residual / update / Poisson loop, which refines to snes_atol),
    timed like the FP64 run, with two references beside it: FP64 with the same
    linear tolerance (separates the precision from the tolerance) and the
    KSP-refined mixed form (each solve refined in FP64 to ksp_rtol).  Both the
    FP64 and the mixed run then continue untimed to --compare-steps, where
    f, phi and the invariant drift are compared."""
    import numpy as np
    import torch
    mrun, melapsed, mrecs, mk = timed_run(w, args, "mixed", args.mixed_kernel)
    mval = args.steps / melapsed
    solves = [p for r in mrecs for p in r["picard"]]
    fp32_solves = [p for p in solves if not p.get("fp64")]
    out = {
        "value": round(mval, 3), "unit": "timesteps/s", "ms_per_step": round(1e3 * melapsed / args.steps, 3),
        "speedup_vs_fp64": round(mval / value64, 3),
        "dtype": "f32 GMRES/P^-1/J.u (f64 dots, Hessenberg) for the Picard solves that start below "
                 "snes_atol * 2^20 (the larger corrections in f64 at the same rtol), inside the f64 Picard "
                 "residual, update and phi",
        "refine": "newton", "fp64_above": 1e-10 * 2.0 ** 20, "inner_rtol": args.inner_rtol, "snes_atol": 1e-10,
        "iterations": dict(_its(mrecs), fp32_solves=len(fp32_solves), fp64_solves=len(solves) - len(fp32_solves),
                           cycles_per_fp32_solve=float(np.mean([p["cycles"] for p in fp32_solves]))
                           if fp32_solves else 0.0),
        "kernels": {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"],
                        "GB/s": round(v["gbs"], 1), "ms": round(v["ms"], 3)} for k, v in mk.items()},
    }
    if not args.no_mixed_refs:
        srun, sel, srecs, _ = timed_run(w, args, "double", None, ksp_rtol=args.inner_rtol)
        sval = args.steps / sel
        out["fp64_same_linear_rtol"] = {"value": round(sval, 3), "ksp_rtol": args.inner_rtol,
                                        "iterations": _its(srecs)}
        out["ratio_vs_fp64_same_linear_rtol"] = round(mval / sval, 3)
        del srun
        arun, ael, arecs, _ = timed_run(w, args, "mixed", None, fp64_above=None)
        out["all_solves_fp32"] = {"value": round(args.steps / ael, 3), "iterations": _its(arecs)}
        del arun
        krun, kel, krecs, _ = timed_run(w, args, "mixed", None, refine="ksp")
        kcy = [p["cycles"] for r in krecs for p in r["picard"]]
        out["ksp_refined"] = {"value": round(args.steps / kel, 3), "speedup_vs_fp64": round(args.steps / kel / value64, 3),
                              "ksp_rtol": 1e-10, "iterations": dict(_its(krecs), cycles_per_solve=float(np.mean(kcy)))}
        del krun
        torch.cuda.empty_cache()
    # continue both runs (untimed) to the comparison step
    done = args.warmup + args.repeats * args.steps
    for _ in range(max(0, args.compare_steps - done)):
        run64.step()
        mrun.step()
    torch.cuda.synchronize()
    n = max(done, args.compare_steps)
    phid = run64.pc_int - run64.pc_int.mean()
    phim = mrun.pc_int - mrun.pc_int.mean()
    i64, imx = run64.invariants(), mrun.invariants()
    drift = lambda r, i: {k: (i[k] - r.inv0[k]) / abs(r.inv0[k]) for k in ("mass", "l2", "energy")}
    out["compare"] = {
        "steps": n,
        "linf_f_vs_fp64": float((mrun.fc - run64.fc).abs().max() / run64.fc.abs().max()),
        "linf_phi_vs_fp64": float((phim - phid).abs().max() / max(phid.abs().max().item(), 1e-300)),
        "drift_fp64": drift(run64, i64), "drift_mixed": drift(mrun, imx),
        "picard_per_step_fp64": float(np.mean([len(r["picard"]) for r in run64.history])),
        "picard_per_step_mixed": float(np.mean([len(r["picard"]) for r in mrun.history])),
        "reverted_fp64": sum(1 for r in run64.history if r.get("reverted")),
        "reverted_mixed": sum(1 for r in mrun.history if r.get("reverted")),
    }
    del mrun
    torch.cuda.empty_cache()
    return out


def adjoint_figures(args):
    """The adjoint workload (bump-on-tail 256x512 J4, dt = 0.05, K steps,
    checkpoint every 20): loss = field energy of the final state, gradient in
    the amplitudes a_m of cos(m k x) M(v), m = 1..16, a_m ~ N(0, 1e-3) from
    default_rng(0), by the discrete adjoint behind torch.autograd.Function;
    checked against central differences (h = 1e-6) along random directions of
    a.  Tolerances are tightened (snes_atol 1e-14, ksp_rtol 1e-13) so that the
    forward map is smooth to well below the difference step."""
    import numpy as np
    import torch
    from vivlasov1d_amd.adjoint import FieldEnergy, ImplicitTimesteps, Stepper, perturbation_basis
    from vivlasov1d_amd.core import Config, bump_on_tail, initial_distribution, normalise
    nx, nv, k, K = 256, 512, 0.3, args.adjoint_steps
    cfg = Config({"grid": {"nx": nx, "nv": nv, "L": 2 * np.pi / k, "vmax": 10.0, "ht": 0.05, "nt": K},
                  "solver": {"poisson_bracket": "arakawaj4", "laplace_operator": "CFD4",
                             "preconditioner_type": "tensor", "preconditioner_scheme": "tensor_kinetic",
                             "petsc_ksp_rtol": 1e-13, "petsc_snes_atol": 1e-14, "petsc_snes_max_iter": 30},
                  "initial_data": {"charge": -1.0}})
    st = Stepper(cfg, ksp_rtol_adj=1e-13, checkpoint=args.adjoint_checkpoint)
    base = torch.from_numpy(normalise(initial_distribution(st.grid, bump_on_tail(k, 0.0)), st.grid)).cuda()
    basis = torch.from_numpy(perturbation_basis(st.grid, k, 16)).cuda()
    rng = np.random.default_rng(0)
    a0 = torch.from_numpy(rng.normal(0.0, 1e-3, 16)).cuda()

    def energy(a):
        return FieldEnergy.apply(ImplicitTimesteps.apply(base + (a[:, None, None] * basis).sum(0), st, K), st)

    a = a0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E = energy(a)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    E.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    grad = a.grad.detach().clone()
    its = list(st.adjoint_its)
    h = 1e-6
    errs = []
    with torch.no_grad():
        for _ in range(args.adjoint_fd_dirs):
            d = torch.from_numpy(rng.standard_normal(16)).cuda()
            d /= d.norm()
            fd = (energy(a0 + h * d).item() - energy(a0 - h * d).item()) / (2 * h)
            got = float((grad * d).sum())
            errs.append(abs(fd - got) / abs(fd))
    del st
    torch.cuda.empty_cache()
    return {"workload": "bump-on-tail k=0.3 256x512 J4 dt=0.05, %d steps, checkpoint %d, 16 modes a_m~N(0,1e-3)"
                        % (K, args.adjoint_checkpoint),
            "loss": E.item(), "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
            "adjoint_gmres_per_step": float(np.mean(its)) if its else None,
            "grad_norm": float(grad.norm()), "fd_h": h, "fd_rel_err": errs,
            "tolerances": {"snes_atol": 1e-14, "ksp_rtol": 1e-13, "adjoint_ksp_rtol": 1e-13}}


def batched_figures(args):
    """Aggregate timesteps/s of the uneven-grid ensemble, batched and sequential."""
    import numpy as np
    import torch
    from vivlasov1d_amd.batch import BatchRun, make_instances, run_sequential
    specs = make_instances(args.batched_instances, seed=0)
    t_setup = time.perf_counter()
    br = BatchRun(specs, graph=True, autotune=True)
    for _ in range(args.batched_warmup):
        br.step()
    br.synchronize()
    t_setup = time.perf_counter() - t_setup
    t0 = time.perf_counter()
    for _ in range(args.batched_steps):
        br.step()
    br.synchronize()
    el = time.perf_counter() - t0
    n = len(specs)
    recs = [r for b in br.buckets for r in b.history[args.batched_warmup:]]
    picard = np.concatenate([r["picard"] for r in recs])
    gm = np.concatenate([r["gmres"] for r in recs])
    out = {
        "instances": n, "buckets": len(br.buckets), "steps": args.batched_steps, "warmup": args.batched_warmup,
        "value": round(n * args.batched_steps / el, 2), "unit": "timesteps/s (all instances)",
        "ms_per_step": round(1e3 * el / args.batched_steps, 3), "setup_s": round(t_setup, 2),
        "points": int(sum(s["nx"] * s["nv"] for s in specs)),
        "picard_per_step": float(picard.mean()), "gmres_per_step": float(gm.mean()),
        "tuning": {"%dx%d" % (b.grid.nx, b.grid.nv): list(b.tuning) for b in br.buckets},
        "graph": True,
    }
    if not args.no_sequential:
        total = args.batched_warmup + args.batched_steps
        el_seq, states = run_sequential(specs, total)
        err = 0.0
        for s_ in specs:
            f, _ = br.state(s_["index"])
            fs, _ = states[s_["index"]]
            err = max(err, float(np.abs(f - fs).max() / np.abs(fs).max()))
        seq_rate = n * total / el_seq
        out["sequential"] = {"value": round(seq_rate, 2), "unit": "timesteps/s (all instances)",
                             "steps": total, "seconds": round(el_seq, 2)}
        out["speedup_vs_sequential"] = round(out["value"] / seq_rate, 2)
        out["linf_f_vs_sequential"] = err
    del br
    torch.cuda.empty_cache()
    return out


def progress(msg):
    """One line per finished leg on stderr (stdout carries only the JSON line)."""
    print("bench: %s (%.0f s)" % (msg, time.perf_counter() - T_START), file=sys.stderr, flush=True)


T_START = time.perf_counter()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--repeats", type=int, default=3,
                    help="timed windows of --steps steps each; value is the median window's rate")
    ap.add_argument("--workload", default="twostream_1024x2048_J4")
    ap.add_argument("--kernel", default="irfft_mdot",
                    help="kernels timed with HIP events in the timed FP64 run ('all', 'none' or a comma list of "
                         "%s); the roofline is the one with the largest total time.  Timed dispatches cost "
                         "throughput (about 4 percent when every launch of the dominant kernel is timed), so the "
                         "default times the dominant kernel only, one launch in --kernel-every" % (KERNELS,))
    ap.add_argument("--kernel-every", type=int, default=8,
                    help="time one launch in every N of --kernel in the timed run (a timed dispatch idles the "
                         "GPU for several microseconds around it; a sample keeps the value run close to an "
                         "untimed one)")
    ap.add_argument("--bracket-kernels", default="jac_rfft,residual,tridiag,maxpy_norm",
                    help="kernels timed in a second, untimed FP64 run: the bracket HBM GB/s (jac_rfft, "
                         "residual) and the other kernels of the step ('none' skips)")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="timesteps of the CPU baseline sample (about 10-20 s on 16 host threads)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="OpenMP threads of the CPU baseline (0: all)")
    ap.add_argument("--cpu1-steps", type=int, default=1,
                    help="timesteps of the one-core CPU sample (0 skips it)")
    ap.add_argument("--no-fusion", action="store_true",
                    help="run the Arnoldi step's irFFT and Gram-Schmidt dots as separate launches")
    ap.add_argument("--no-config2-bracket", action="store_true",
                    help="skip the bracket GB/s at the 256x512 nonlinear-Landau size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision timing")
    ap.add_argument("--mixed-kernel", default="none", help="kernel timed in the mixed run (e.g. mdot32)")
    ap.add_argument("--inner-rtol", type=float, default=1e-5, help="FP32 cycle reduction of the mixed form")
    ap.add_argument("--no-mixed-refs", action="store_true",
                    help="skip the FP64-at-the-same-tolerance and KSP-refined mixed references")
    ap.add_argument("--compare-steps", type=int, default=1000,
                    help="timestep at which the mixed and FP64 runs are compared (f, phi, invariant drift)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-checkpoint", type=int, default=20)
    ap.add_argument("--adjoint-fd-dirs", type=int, default=2, help="finite-difference directions of the check")
    ap.add_argument("--no-sequential", action="store_true", help="skip the sequential run of the ensemble")
    ap.add_argument("--batched-instances", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=3)
    ap.add_argument("--batched-warmup", type=int, default=1)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (reference kernel for the HBM byte counters)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n,
                          "copies": 5}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the timestep is not sharded across GPUs "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    w = workloads()[args.workload]
    N = w["nx"] * w["nv"]
    run, elapsed, recs, kstats = timed_run(w, args, "double", args.kernel, every=args.kernel_every)
    progress("FP64 timed")
    dom = max(kstats, key=lambda k: kstats[k]["ms"]) if kstats else None
    roof = None
    if dom:
        d = kstats[dom]
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": d["bytes"] / d["launches"], "avg_launch_us": round(d["avg_us"], 3),
                "launches": d["launches"], "sampled_every": args.kernel_every}
        # HBM bytes per launch measured by the PMC runs summarised in profiles/
        from bench_profiles import read_summary
        src = os.path.join("profiles", "%s_%s.txt" % (dom, args.workload))
        tk = read_summary(os.path.join(os.path.dirname(os.path.abspath(__file__)), src))
        if "traffic_bytes_per_launch" in tk:
            roof["traffic"] = tk["traffic_bytes_per_launch"]
            roof["traffic_source"] = src
            # counter bytes over the algorithmic bytes of the same dispatches (the
            # PMC runs' timed region), as the summary records it
            if "traffic_over_algorithmic" in tk:
                roof["traffic_over_algorithmic"] = tk["traffic_over_algorithmic"]
    bracket = step_kernels = None
    if args.bracket_kernels not in ("", "none"):
        brun, _, _, bk = timed_run(w, args, "double", args.bracket_kernels, repeats=1)
        progress("bracket kernels timed")
        del brun
        step_kernels = {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"], "GB/s": round(v["gbs"], 1),
                            "frac_hbm_peak": round(v["gbs"] / HBM_PEAK_GBS, 4),
                            "bytes_per_launch": v["bytes"] / v["launches"]} for k, v in bk.items()}
        bracket = {k: v for k, v in step_kernels.items() if k in ("jac_rfft", "jacobian", "residual")}
    sb = sum(step_bytes(N, r) for r in recs)
    picard = [len(r["picard"]) for r in recs]
    gm = [p["gmres"] for r in recs for p in r["picard"]]

    cpu = cpu1 = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(w, args.cpu_steps, args.cpu_threads)
            progress("CPU baseline (all cores)")
            if args.cpu1_steps > 0:
                cpu1 = cpu_baseline(w, args.cpu1_steps, 1)
                progress("CPU baseline (1 core)")
        except Exception as e:      # reported, never silently replaced
            cpu = {"error": "%s: %s" % (type(e).__name__, e)}
    bracket2 = None if args.no_config2_bracket else bracket_config2(args)
    progress("bracket at 256x512")

    inv = run.invariants()
    value = args.steps / elapsed
    run_windows = run.windows

    mixed = None
    if not args.no_mixed:
        mixed = mixed_figures(w, args, run, value)
        progress("mixed precision")
        del run
        torch.cuda.empty_cache()
    batched = None
    if not args.no_batched:
        batched = batched_figures(args)
        progress("batched ensemble")
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_figures(args)
        progress("adjoint")

    line = {
        "metric": "timesteps/s at 1024x2048 J4 (plain FP64)" if args.workload.startswith("twostream_1024")
        else "timesteps/s %s" % args.workload,
        "value": round(value, 3), "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic initial condition (%s), generated in-process" % w["desc"],
        "config": {"workload": args.workload, "nx": w["nx"], "nv": w["nv"], "bracket": w["bracket"],
                   "dt": w["ht"], "ksp_rtol": 1e-10, "snes_atol": 1e-10, "restart": 30,
                   "preconditioner": "tensor_kinetic", "parallelism": "single GPU"},
        "roofline": roof,
        "step_roofline": {"bytes_per_step": sb / max(1, len(recs)),
                          "achieved": round(sb / max(1, len(recs)) * args.steps / elapsed / 1e9, 1), "unit": "GB/s",
                          "frac": round(sb / max(1, len(recs)) * args.steps / elapsed / 1e9 / HBM_PEAK_GBS, 4)},
        "iterations": {"picard_per_step": float(np.mean(picard)), "gmres_per_solve": float(np.mean(gm)),
                       "gmres_per_step": float(sum(gm)) / max(1, len(recs))},
        "kernels": {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"], "GB/s": round(v["gbs"], 1),
                        "ms": round(v["ms"], 3)} for k, v in kstats.items()},
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "timing": {"repeats": args.repeats, "window_s": [round(t, 4) for t in run_windows],
                   "value_from": "median window"},
        "invariants": {k: inv[k] for k in ("mass", "l2", "energy")},
        "bracket": bracket,
        "step_kernels": step_kernels,
        "bracket_landau_256x512": bracket2,
        "mixed": mixed,
        "batched": batched,
        "adjoint": adjoint,
    }
    if cpu and "value" in cpu and cpu["value"]:
        line["gpu_over_cpu"] = round(value / cpu["value"], 2)
    if cpu1 and "value" in cpu1 and cpu1["value"]:
        line["gpu_over_cpu_1core"] = round(value / cpu1["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
